// handles.h -- definitions of the opaque handle types of the C interface.
// The reference's versions hold host pointers (qphix_d3_internal.h:7-32); these
// own device memory.
#pragma once
#include "common.h"

#include <functional>

struct QPHIX_D3_ColorVector_struct : qphix::VectorT<double> {};
struct QPHIX_F3_ColorVector_struct : qphix::VectorT<float> {};
struct QPHIX_D3_FermionLinksAsqtad_struct : qphix::LinksT<double> {};
struct QPHIX_F3_FermionLinksAsqtad_struct : qphix::LinksT<float> {};
struct QPHIX_D3_GaugeField_struct : qphix::GaugeT<double> {};
struct QPHIX_F3_GaugeField_struct : qphix::GaugeT<float> {};
struct QPHIX_D3_Force_struct : qphix::ForceT<double> {};
struct QPHIX_F3_Force_struct : qphix::ForceT<float> {};

namespace qphix {
template <typename T> struct Handles;
template <> struct Handles<double> {
  using V = QPHIX_D3_ColorVector;
  using L = QPHIX_D3_FermionLinksAsqtad;
  using G = QPHIX_D3_GaugeField;
  using F = QPHIX_D3_Force;
};
template <> struct Handles<float> {
  using V = QPHIX_F3_ColorVector;
  using L = QPHIX_F3_FermionLinksAsqtad;
  using G = QPHIX_F3_GaugeField;
  using F = QPHIX_F3_Force;
};

template <typename T> void links_free(LinksT<T> *L);
// A link handle built by the import pipeline of asqtad_create_L_from_raw from
// raw MILC-order blocks that fill(cb, stage) writes on the device (fat, fat
// back, long, long back at stage + k * 72 vh); null (error recorded) when
// memory runs out or fill fails (fields.hip).
template <typename T>
typename Handles<T>::L *create_L_device(int evenodd, const std::function<bool(int, T *)> &fill);
int long_recon_mode();
} // namespace qphix

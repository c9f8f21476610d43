"""torch.autograd harness of the HISQ/asqtad fermion force.

FermionAction.apply(fat, lng, session, phi, masses, coef, resid) returns
    S = sum_k coef[k] Re phi^dag (4 masses[k]^2 - D_eo D_oe)^{-1} phi
for forward fat and long links given as complex128 tensors (V, 4, 3, 3) in
raw MILC site order, KS phases folded in (MILC's back links -U^dag(x - mu) are
derived from them).  The forward pass is the FP64 multi-shift solver
(QPHIX_D3_asqtad_invert_multi); the backward pass is the adjoint
outer-product kernel (QPHIX_D3_asqtad_fermion_force_oprod), whose outputs O
give PyTorch's gradient of a real function of complex leaves,
dS/dRe U + i dS/dIm U = 2 O^dag.  Both passes run in libqphixmilc_hip.so; the
tensors are only the carriers.
"""
import numpy as np
import torch

from . import api as A
from . import fields as F


class _Solved:
    """Device links and solutions kept between forward and backward."""

    def __init__(self, L, xs, vb):
        self.L, self.xs, self.vb = L, xs, vb

    def free(self):
        for o in [self.L, self.vb] + list(self.xs or []):
            if o is not None:
                o.free()
        self.L, self.xs, self.vb = None, None, None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def solve(session, fat, lng, phi_raw, masses, resid, niter=4000, nrestart=5):
    """Device links from forward links, and the multi-shift solutions."""
    lat = F.Lattice(session.dims)
    L = A.Links(session, *F.raw_link_set(lat, fat, lng))
    vb = A.Vector(session, phi_raw, A.EVEN)
    xs = [A.Vector(session, None, A.EVEN) for _ in masses]
    ret, info, res = A.invert_multi(L, xs, vb, list(masses), A.EVEN, resid, niter=niter,
                                    nrestart=nrestart)
    if info.status != 0:
        _Solved(L, xs, vb).free()
        raise RuntimeError("multi-shift solve did not converge")
    return _Solved(L, xs, vb)


def action_value(phi_raw, xs_raw, coef, vh):
    """sum_k coef[k] Re phi^dag X_k over the even half."""
    p = phi_raw[:6 * vh].view(np.complex128)
    return float(sum(c * np.real(np.vdot(p, x[:6 * vh].view(np.complex128)))
                     for c, x in zip(coef, xs_raw)))


class FermionAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fat, lng, session, phi_raw, masses, coef, resid):
        f = fat.detach().cpu().numpy()
        l = lng.detach().cpu().numpy()
        st = solve(session, f, l, phi_raw, masses, resid)
        S = action_value(np.asarray(phi_raw, dtype=np.float64), [x.extract() for x in st.xs],
                         coef, session.vh)
        ctx.st = st
        ctx.coef = list(coef)
        ctx.device = fat.device
        return torch.tensor(S, dtype=torch.float64, device=fat.device)

    @staticmethod
    def backward(ctx, gS):
        st = ctx.st
        o1, o3, _ = A.fermion_force_oprod(st.L, ctx.coef, st.xs)
        st.free()
        g = float(gS)
        g1 = torch.from_numpy(2.0 * g * np.conj(np.swapaxes(o1, -1, -2))).to(ctx.device)
        g3 = torch.from_numpy(2.0 * g * np.conj(np.swapaxes(o3, -1, -2))).to(ctx.device)
        return g1, g3, None, None, None, None, None

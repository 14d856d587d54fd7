"""The one-call Green's function entry points of the C ABI (ed_greens_cf,
ed_greens_lehmann_t0), called directly as a non-Python binder would.

* ed_greens_cf with both neighbour sectors in one call equals, bit for bit,
  the Python layer's per-part calls (0 + p + h either way), and equals the
  oracle's exact Lehmann G once the fraction spans the sectors (nsteps >= D).
* ed_greens_lehmann_t0 against the oracle's lehmann_t0, both spins, both
  sign modes.
* Sector checks: an operator of the wrong sector is refused.
"""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu


def _setup(ns=8, n_up=4, n_dn=4, u=3.0, seed=3):
    from cmpy_amd import HubbardModel, chain_neighbors, groundstate

    model = HubbardModel(ns, chain_neighbors(ns), inter=u, hop=1.0, mu=u / 2)
    gs = groundstate(model, n_up, n_dn, tol=1e-12, device=True, seed=seed)
    return model, gs


def _g(call, z, *args):
    from cmpy_amd import _lib as L

    zt = torch.from_numpy(np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())).cuda()
    g = torch.empty_like(zt)
    call(zt, g)
    gh = g.view(-1, 2).cpu().numpy()
    return gh[:, 0] + 1j * gh[:, 1]


@pytest.mark.parametrize("sigma", [1, 2])
def test_greens_cf_one_call(gpu, sigma):
    from cmpy_amd import _lib as L
    from cmpy_amd import greens_cf
    from cmpy_amd.operators import SIGN_MODES

    ns, n = 8, 4
    model, gs = _setup(ns, n, n)
    d = (1, 0) if sigma == 1 else (0, 1)
    op_p = model.hamilton_operator(n_up=n + d[0], n_dn=n + d[1])
    op_m = model.hamilton_operator(n_up=n - d[0], n_dn=n - d[1])
    z = np.linspace(-6, 6, 97) + 0.3j
    nsteps = max(op_p.size, op_m.size)
    r = greens_cf(model, z, n, n, sigma=sigma, nsteps=nsteps, gs=gs, reorth="full")
    nbytes = C.c_int64()
    L.call("ed_greens_cf_workspace", op_p.handle, op_m.handle, nsteps, 1, C.byref(nbytes))
    work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    parts = L.CfParts()

    def call(zt, g):
        L.call("ed_greens_cf", op_p.handle, op_m.handle, ns, n, n, 0, sigma, SIGN_MODES["reference"],
               L.dptr(gs.state), gs.energy, nsteps, 1, L.dptr(zt), zt.numel() // 2, L.dptr(g), C.byref(parts),
               L.dptr(work), nbytes.value, L.stream_ptr())

    g = _g(call, z)
    assert np.array_equal(g, r.gf)
    assert list(parts.steps) == [len(r.alpha_p), len(r.alpha_h)]
    assert parts.norm2[0] == r.norm2_p and parts.norm2[1] == r.norm2_h
    # the fraction spans both sectors: exact
    bonds = O.chain_bonds(ns)
    mk = lambda a, b: O.hubbard_operator(ns, a, b, bonds, 3.0, 0.0, 1.0, mu=1.5)
    psi = gs.state.cpu().numpy()
    o_n, o_p, o_m = mk(n, n), mk(n + d[0], n + d[1]), mk(n - d[0], n - d[1])
    cd = O.apply_cdag(psi, ns, n, n, 0, sigma)
    c = O.apply_c(psi, ns, n, n, 0, sigma)
    ref = O.lehmann_t0(o_n, o_p, o_m, psi, gs.energy, cd, c, z)
    assert np.abs(g - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("sigma,mode", [(1, "reference"), (2, "jordan_wigner")])
def test_greens_lehmann_t0_one_call(gpu, sigma, mode):
    from cmpy_amd import _lib as L
    from cmpy_amd.operators import SIGN_MODES

    ns, n = 8, 4
    model, gs = _setup(ns, n, n)
    d = (1, 0) if sigma == 1 else (0, 1)
    op_p = model.hamilton_operator(n_up=n + d[0], n_dn=n + d[1])
    op_m = model.hamilton_operator(n_up=n - d[0], n_dn=n - d[1])
    z = np.linspace(-6, 6, 61) + 0.2j
    nbytes = C.c_int64()
    L.call("ed_greens_lehmann_t0_workspace", op_p.handle, op_m.handle, C.byref(nbytes))
    work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")

    def call(zt, g):
        L.call("ed_greens_lehmann_t0", op_p.handle, op_m.handle, ns, n, n, 2, sigma, SIGN_MODES[mode],
               L.dptr(gs.state), gs.energy, L.dptr(zt), zt.numel() // 2, L.dptr(g), L.dptr(work), nbytes.value,
               L.stream_ptr())

    g = _g(call, z)
    bonds = O.chain_bonds(ns)
    mk = lambda a, b: O.hubbard_operator(ns, a, b, bonds, 3.0, 0.0, 1.0, mu=1.5)
    psi = gs.state.cpu().numpy()
    cd = O.apply_cdag(psi, ns, n, n, 2, sigma, mode)
    c = O.apply_c(psi, ns, n, n, 2, sigma, mode)
    ref = O.lehmann_t0(mk(n, n), mk(n + d[0], n + d[1]), mk(n - d[0], n - d[1]), psi, gs.energy, cd, c, z)
    assert np.abs(g - ref).max() <= 1e-11 * np.abs(ref).max()


def test_greens_one_call_sector_checks(gpu):
    from cmpy_amd import _lib as L

    ns, n = 8, 4
    model, gs = _setup(ns, n, n)
    op_wrong = model.hamilton_operator(n_up=n, n_dn=n)
    op_m = model.hamilton_operator(n_up=n - 1, n_dn=n)
    nbytes = C.c_int64()
    with pytest.raises(L.EDInvalidArgument, match="op_p holds sector"):
        L.call("ed_greens_lehmann_t0", op_wrong.handle, op_m.handle, ns, n, n, 0, 1, 0, L.dptr(gs.state),
               gs.energy, None, 0, None, None, 0, L.stream_ptr())
    L.call("ed_greens_cf_workspace", None, op_m.handle, 10, 0, C.byref(nbytes))
    assert nbytes.value > 0

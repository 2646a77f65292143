"""Batched sweep (one graph launch over many uneven systems) on the MI355X.

Each system of the batch must give what the single-system path gives for it:
the FP64 batch matches the FP64 oracle to 1e-10 (eigenvalues and spectrum),
the FP32-storage batch with FP64 refinement matches the oracle's lowest
excitons to 1e-10 and its spectrum to the FP32-storage level (1e-5 relative
L2), and repeated launches of the same graph are bitwise identical.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
import bse_oracle as O

pytestmark = pytest.mark.gpu

N_LOW = 8


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


def mixed_batch():
    """Uneven 1v+1c honeycomb systems plus a 2v+2c spinful one (MFMA assembly)."""
    sweep = bse_amd.honeycomb_sweep(5, seed=11, n_min=4, n_max=9)
    return sweep + [bse_amd.honeycomb_magnetic(4)]


def _oracle(systems):
    out = []
    for s in systems:
        p = O.problem_from_system(s)
        out.append((p, O.solve(p)))
    return out


def _omegas(p):
    return np.asarray(p.omegas)


def _ref_eps(p, sol, om):
    return O.dielectric(sol["E"], sol["f"], om, p.eta, O.C_EV_A / (p.volume * p.nk))


def test_batch_fp64_matches_oracle(ctx):
    systems = mixed_batch()
    refs = _oracle(systems)
    om = _omegas(refs[0][0])
    print("batch fp64 case: sizes", [s.n_bse for s in systems], flush=True)
    ctx.batch_prepare(systems, precision="fp64", n_lowest=N_LOW, refine=False, omegas=om,
                      eta=refs[0][0].eta)
    tot, per = ctx.batch_run()
    ev, eps = ctx.batch_results()
    assert tot > 0 and np.all(per > 0) and np.all(per <= tot * 1.01 + 0.05)
    for i, (p, sol) in enumerate(refs):
        scale = np.abs(sol["E"]).max()
        assert np.abs(ev[i] - sol["E"][:N_LOW]).max() < 1e-10 * scale, i
        ref_eps = _ref_eps(p, sol, om)
        rel = np.linalg.norm(eps[i].imag - ref_eps.imag) / np.linalg.norm(ref_eps.imag)
        assert rel < 1e-10, (i, rel)
        rel_re = np.linalg.norm(eps[i].real - ref_eps.real) / np.linalg.norm(ref_eps.real)
        assert rel_re < 1e-10, (i, rel_re)
    info = ctx.batch_info()
    assert info["systems"] == len(systems) and info["graph_nodes"] > 0
    assert info["max_n"] == max(s.n_bse for s in systems)


@pytest.mark.parametrize("prec", ["fp32", "bf16x2"])
def test_batch_mixed_refined_matches_oracle(ctx, prec):
    systems = mixed_batch()
    refs = _oracle(systems)
    om = _omegas(refs[0][0])
    print(f"batch {prec} refined case", flush=True)
    ctx.batch_prepare(systems, precision=prec, n_lowest=N_LOW, refine=True, omegas=om,
                      eta=refs[0][0].eta)
    ctx.batch_run()
    ev, eps = ctx.batch_results()
    for i, (p, sol) in enumerate(refs):
        scale = np.abs(sol["E"]).max()
        assert np.abs(ev[i] - sol["E"][:N_LOW]).max() < 1e-10 * scale, i
        ref_eps = _ref_eps(p, sol, om)
        rel = np.linalg.norm(eps[i].imag - ref_eps.imag) / np.linalg.norm(ref_eps.imag)
        assert rel < (1e-5 if prec == "fp32" else 1e-3), (i, rel)


def test_batch_matches_single_system_path(ctx):
    systems = bse_amd.honeycomb_sweep(4, seed=3, n_min=5, n_max=8)
    p0 = O.problem_from_system(systems[0])
    om = _omegas(p0)
    ctx.batch_prepare(systems, precision="fp32", n_lowest=N_LOW, refine=True, omegas=om,
                      eta=p0.eta)
    ctx.batch_run()
    ev, eps = ctx.batch_results()
    for i, s in enumerate(systems):
        ctx.set_system(s).build("fp32").eigh(N_LOW, refine=True)
        E = ctx.eigenvalues()
        e1 = ctx.spectrum(om, p0.eta)
        assert np.abs(ev[i] - E[:N_LOW]).max() < 1e-10 * np.abs(E).max()
        assert np.linalg.norm(eps[i] - e1) / np.linalg.norm(e1) < 1e-6


def test_batch_relaunch_is_bitwise_stable(ctx):
    systems = bse_amd.honeycomb_sweep(6, seed=5, n_min=4, n_max=10)
    p0 = O.problem_from_system(systems[0])
    ctx.batch_prepare(systems, precision="fp32", n_lowest=4, refine=True, omegas=_omegas(p0),
                      eta=p0.eta)
    ctx.batch_run()
    ev1, eps1 = ctx.batch_results()
    ctx.batch_run()
    ev2, eps2 = ctx.batch_results()
    assert np.array_equal(ev1, ev2)
    assert np.array_equal(eps1, eps2)


def test_batch_rejects_oversized_system(ctx):
    big = bse_amd.random_multiorbital(17)    # N = 16 * 289 = 4624 > 4096
    with pytest.raises(_abi.BSEError):
        ctx.batch_prepare([big], precision="fp32", n_lowest=4)

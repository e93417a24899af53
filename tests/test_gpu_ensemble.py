"""Ensembles of uneven-N lattices in one persistent launch per step.

Every realisation must equal a single-lattice run of the same steps bit for bit
(so the output does not depend on the batch composition), graph replays must
equal the eager run (replayed several times with eager work in between), and
the autotuner's plan must not change any bit. Precisions: FP64 (0), FP32-comp (1)
and the compensated storage with float forces (2, CF_PREC_F32S, paced or not).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_parity import assert_bitwise, host

pytestmark = pytest.mark.gpu

FRW = dict(a1=1.0, H1=0.5, t1=1.0)
NS = [16, 24, 32, 16, 40, 24]


def raw_state(cfl, ens, r):
    """hi, lo (F32C, F32S) or x (F64) of realisation r as fresh torch tensors."""
    N = ens.Ns[r]
    n = 2 * N ** 3
    x, lo = ens.state_ptrs(r)
    if ens.precision != cfl.CF_PREC_F64:
        hi_t = torch.empty(n, dtype=torch.float32, device="cuda")
        lo_t = torch.empty_like(hi_t)
        cfl.check(cfl.lib.cf_d2d(cfl._ptr(hi_t), C.c_void_p(x), n * 4, cfl._stream()))
        cfl.check(cfl.lib.cf_d2d(cfl._ptr(lo_t), C.c_void_p(lo), n * 4, cfl._stream()))
        return hi_t, lo_t
    x_t = torch.empty(n, dtype=torch.float64, device="cuda")
    cfl.check(cfl.lib.cf_d2d(cfl._ptr(x_t), C.c_void_p(x), n * 8, cfl._stream()))
    return (x_t,)


def make(cfl, prec, lam=0.1):
    ens = cfl.Ensemble(NS, [1.0 * n for n in NS], 1.0, lam, cfl.radiation(**FRW), prec)
    ens.init_grf(list(range(100, 100 + len(NS))), cfl.CF_SPEC_GAUSS, (0.5,), 0.5)
    return ens


def check_single_lattice_runs(cfl, ens, prec, init, t0, k0, steps, dt, lam=0.1):
    """Every realisation of `ens` equals the single-lattice run from its initial state."""
    for r, N in enumerate(ens.Ns):
        p = cfl.kg_params(N, float(N), 1.0, lam, cfl.radiation(**FRW))
        if prec != cfl.CF_PREC_F64:
            hi, lo = init[r]
            if prec == cfl.CF_PREC_F32C:
                cfl.rk4_run_f32c(p, hi, lo, t0, k0, steps, dt, cfl.alloc_work(N, prec))
            else:
                cfl.rk4_run_tb_f32s(p, hi, lo, t0, k0, steps, dt, cfl.alloc_tb_work(N, prec))
            ghi, glo = raw_state(cfl, ens, r)
            assert_bitwise(host(ghi), host(hi), f"hi r={r}")
            assert_bitwise(host(glo), host(lo), f"lo r={r}")
        else:
            (x,) = init[r]
            cfl.rk4_run_f64(p, x, t0, k0, steps, dt, cfl.alloc_work(N, prec))
            (gx,) = raw_state(cfl, ens, r)
            assert_bitwise(host(gx), host(x), f"x r={r}")


@pytest.mark.parametrize("prec,grid", [(0, 0), (1, 0), (0, -1), (1, -1), (2, 0), (2, -1)])
def test_ensemble_equals_single_lattice_runs(cfl, prec, grid):
    """Persistent (default grid) and flat (grid = CF_ENS_GRID_FLAT) launches."""
    print(f"case ensemble prec={prec} grid={grid}", flush=True)
    ens = make(cfl, prec)
    if grid:
        ens.set_plan(0, 0, grid)
        assert ens.get_plan(NS[0])[1] == grid
    t0, k0, steps, dt = 1.0, 2, 3, 0.05
    init = [raw_state(cfl, ens, r) for r in range(ens.n)]
    ens.run(t0, k0, steps, dt)
    check_single_lattice_runs(cfl, ens, prec, init, t0, k0, steps, dt)
    ens.close()


@pytest.mark.parametrize("prec,grid", [(0, 0), (1, 0), (0, -1), (1, -1), (2, 0), (2, -1)])
def test_region_shapes_keep_bits(cfl, prec, grid):
    """Blocks over 24 x 40 regions (16 x 32 tiles: one partial tile per side at N=24
    and N=40, a single tile at N=16) mixed with 32 x 32 ones in the same launches
    give every realisation the bits of its single-lattice run."""
    print(f"case ensemble regions prec={prec} grid={grid}", flush=True)
    ens = make(cfl, prec)
    if grid:
        ens.set_plan(0, 0, grid)
    for n in (16, 40, 24):
        ens.set_region(n, cfl.CF_ENS_REGION_24X40)
        assert ens.get_region(n) == cfl.CF_ENS_REGION_24X40
    assert ens.get_region(32) == cfl.CF_ENS_REGION_32X32
    ens.set_plan(40, 16, 0)  # three chunks along a, the last one partial
    t0, k0, steps, dt = 1.0, 2, 3, 0.05
    init = [raw_state(cfl, ens, r) for r in range(ens.n)]
    ens.run(t0, k0, steps, dt)
    check_single_lattice_runs(cfl, ens, prec, init, t0, k0, steps, dt)
    ens.close()


@pytest.mark.parametrize("prec,grid", [(0, 0), (1, 0), (0, -1), (1, -1), (2, 0), (2, -1)])
def test_ensemble_without_lambda(cfl, prec, grid):
    """lambda = 0 selects the kernels without the phi^3 term; both region shapes."""
    print(f"case ensemble lambda=0 prec={prec} grid={grid}", flush=True)
    ens = make(cfl, prec, lam=0.0)
    if grid:
        ens.set_plan(0, 0, grid)
    ens.set_region(40, cfl.CF_ENS_REGION_24X40)
    t0, k0, steps, dt = 1.0, 0, 2, 0.05
    init = [raw_state(cfl, ens, r) for r in range(ens.n)]
    ens.run(t0, k0, steps, dt)
    check_single_lattice_runs(cfl, ens, prec, init, t0, k0, steps, dt, lam=0.0)
    ens.close()


def test_ensemble_oracle_small(cfl, O):
    """One realisation against the oracle, from the same initial state."""
    ens = make(cfl, 0)
    x0 = host(ens.state(1))
    ens.run(1.0, 0, 2, 0.05)
    eq = O.Equation(24, 24.0, 1.0, 0.1, O.Cosmo.radiation(**FRW))
    x = x0.copy()
    eq.rk4_run(x, 1.0, 0.05, 2)
    assert_bitwise(host(ens.state(1)), x, "ensemble r=1 vs oracle")
    ens.close()


@pytest.mark.parametrize("prec,grid", [(0, 0), (1, 0), (1, -1), (2, -1)])
def test_graph_replays_equal_eager(cfl, prec, grid):
    print(f"case graph prec={prec} grid={grid}", flush=True)
    seeds = list(range(100, 100 + len(NS)))
    ens = make(cfl, prec)
    if grid:
        ens.set_plan(0, 0, grid)
    ens.run(1.0, 0, 4, 0.05, use_graph=False)
    torch.cuda.synchronize()
    eager = [host(ens.state(r)) for r in range(ens.n)]
    for rep in range(3):
        # eager work between replays: re-initialise (GRF + FFT + rescale) and sync
        ens.init_grf(seeds, cfl.CF_SPEC_GAUSS, (0.5,), 0.5)
        torch.cuda.synchronize()
        ens.run(1.0, 0, 4, 0.05, use_graph=True)
        torch.cuda.synchronize()
        for r in range(ens.n):
            assert_bitwise(host(ens.state(r)), eager[r], f"replay {rep} r={r}")
    ens.close()


@pytest.mark.parametrize("prec", [1, 2])
def test_autotuner_keeps_bits(cfl, prec):
    print(f"case ensemble tune prec={prec}", flush=True)
    ens = make(cfl, prec)
    x0 = [raw_state(cfl, ens, r) for r in range(ens.n)]
    g = ens.tune(0.05, reps=2)
    assert g > 0 or g == cfl.CF_ENS_GRID_FLAT
    # tuning did not touch the state
    for r in range(ens.n):
        for a, b in zip(raw_state(cfl, ens, r), x0[r]):
            assert torch.equal(a, b)
    print(f"tuned grid {g} pace {ens.get_pace()}", flush=True)
    ens.run(1.0, 0, 2, 0.05)
    after_tuned = [host(ens.state(r)) for r in range(ens.n)]
    ens.close()
    ens2 = make(cfl, prec)
    ens2.set_plan(0, 0, 97)  # an arbitrary different plan
    ens2.run(1.0, 0, 2, 0.05)
    for r in range(ens2.n):
        assert_bitwise(host(ens2.state(r)), after_tuned[r], f"plan-independence r={r}")
    ens2.close()


def test_density_spectrum(cfl, O):
    ens = make(cfl, 1)
    ens.run(1.0, 0, 2, 0.05)
    t = 1.1
    for r in (0, 4):
        N = NS[r]
        x = host(ens.state(r))
        eq = O.Equation(N, float(N), 1.0, 0.1, O.Cosmo.radiation(**FRW))
        ref = O.power_spectrum(N, eq.energy_density(x, t))
        g = host(ens.density_spectrum(r, t))
        assert np.all(g[ref == 0] == 0)
        assert np.all(np.abs(g - ref) <= 1e-10 * ref + 1e-18 * ref.max())
    ens.close()


def test_flat_grouped_xcd_order(cfl):
    """8 lattices of N=32 (2x2 tiles each: 32 work units, a multiple of 8 XCDs x 4)
    take the flat launch's grouped XCD order; every realisation still equals its
    single-lattice run bit for bit (each unit ran exactly once)."""
    Ns = [32] * 8
    ens = cfl.Ensemble(Ns, [32.0] * 8, 1.0, 0.1, cfl.radiation(**FRW), cfl.CF_PREC_F32C)
    ens.init_grf(list(range(300, 308)), cfl.CF_SPEC_GAUSS, (0.5,), 0.5)
    ens.set_plan(0, 0, cfl.CF_ENS_GRID_FLAT)
    for n in set(Ns):
        ens.set_plan(n, n, 0)
    init = [raw_state(cfl, ens, r) for r in range(ens.n)]
    ens.run(1.0, 0, 2, 0.05)
    p = cfl.kg_params(32, 32.0, 1.0, 0.1, cfl.radiation(**FRW))
    for r in range(ens.n):
        hi, lo = init[r]
        cfl.rk4_run_f32c(p, hi, lo, 1.0, 0, 2, 0.05, cfl.alloc_work(32, cfl.CF_PREC_F32C))
        ghi, glo = raw_state(cfl, ens, r)
        assert_bitwise(host(ghi), host(hi), f"hi r={r}")
        assert_bitwise(host(glo), host(lo), f"lo r={r}")
    ens.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_segments_without_host_sync(cfl, use_graph):
    """Two runs of 3 steps issued back to back (no synchronisation in between: the
    second run's step table is uploaded while the first run may still be reading its
    own) equal one run of 6 steps bit for bit."""
    print(f"case segments use_graph={use_graph}", flush=True)
    a = make(cfl, 1)
    b = make(cfl, 1)
    torch.cuda.synchronize()
    a.run(1.0, 0, 3, 0.05, use_graph=use_graph)
    a.run(1.0, 3, 3, 0.05, use_graph=use_graph)
    b.run(1.0, 0, 6, 0.05, use_graph=False)
    torch.cuda.synchronize()
    for r in range(a.n):
        for x, y in zip(raw_state(cfl, a, r), raw_state(cfl, b, r)):
            assert_bitwise(host(x), host(y), f"segments r={r}")
    a.close()
    b.close()


def test_graph_not_replayed_after_eager_run(cfl):
    """A captured graph (steps k0 = 0..3), then an eager run of other steps (k0 = 4..7,
    which rewrites the step table), then the first call again: the last call must run
    the steps of k0 = 0..3, not replay a graph that reads the rewritten table."""
    a = make(cfl, 1)
    b = make(cfl, 1)
    a.run(1.0, 0, 4, 0.05, use_graph=True)
    a.run(1.0, 4, 4, 0.05, use_graph=False)
    a.run(1.0, 0, 4, 0.05, use_graph=True)
    for k0 in (0, 4, 0):
        b.run(1.0, k0, 4, 0.05, use_graph=False)
    torch.cuda.synchronize()
    for r in range(a.n):
        for x, y in zip(raw_state(cfl, a, r), raw_state(cfl, b, r)):
            assert_bitwise(host(x), host(y), f"graph after eager r={r}")
    a.close()
    b.close()


@pytest.mark.parametrize("grid", [0, -1])
def test_paced_f32s_ensemble_keeps_bits(cfl, grid):
    """A paced single-precision-stencil ensemble (both launch forms, graph replay)
    gives every realisation the bits of its single-lattice run; pacing is refused
    for the other precisions."""
    print(f"case ensemble f32s paced grid={grid}", flush=True)
    ens = make(cfl, cfl.CF_PREC_F32S)
    if grid:
        ens.set_plan(0, 0, grid)
    ens.set_region(40, cfl.CF_ENS_REGION_24X40)
    ens.set_pace(180)
    assert ens.get_pace() == 180
    t0, k0, steps, dt = 1.0, 1, 3, 0.05
    init = [raw_state(cfl, ens, r) for r in range(ens.n)]
    ens.run(t0, k0, steps, dt, use_graph=True)
    check_single_lattice_runs(cfl, ens, cfl.CF_PREC_F32S, init, t0, k0, steps, dt)
    ens.close()
    other = make(cfl, cfl.CF_PREC_F32C)
    with pytest.raises(cfl.CfError):
        other.set_pace(100)
    other.set_pace(0)
    other.close()

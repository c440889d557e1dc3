"""The snapshot packers' restatement (oracle/snapshots.py) against the
behaviours the reference's own tests fix (tests/py/test_mpiio_fields.py,
test_mpiio_particles.py, test_mpiio_spectra.py) and closed forms."""

import numpy as np
import pytest

import snapshots as osnap

F32 = np.float32


def grid_arrays(nt, n, fill):
    X = [a * b for a, b in zip(nt, n)]
    return [np.full(X, v, dtype=F32) for v in fill]


def cell_centre_particles(nt, n, ppc=1, nsp=2):
    """parts[t][s]: [3, k] float32 positions, ppc per cell at the cell centre."""
    parts = []
    for t in range(nt[0] * nt[1] * nt[2]):
        tc = (t // (nt[1] * nt[2]), (t // nt[2]) % nt[1], t % nt[2])
        g = np.stack(np.meshgrid(*(np.arange(v) + 0.5 for v in n), indexing="ij"), 0).reshape(3, -1)
        g = g + np.array([tc[a] * n[a] for a in range(3)]).reshape(3, 1)
        p = np.repeat(g, ppc, 1).astype(F32)
        parts.append([p.copy() for _ in range(nsp)])
    return parts


@pytest.mark.parametrize("stride", [1, 2, 4, 8])
def test_uniform_fields_and_density(stride):
    # test_single_tile_field_values, test_stride, test_stride_equals_mesh_size,
    # test_uniform_density_ppc1 / ppc4, test_density_with_stride
    nt, n = (2, 2, 1), (8, 8, 8)
    E = grid_arrays(nt, n, (1, 2, 3))
    B = grid_arrays(nt, n, (4, 5, 6))
    J = grid_arrays(nt, n, (7, 8, 9))
    for ppc in (1, 4):
        out = osnap.fields_pack(E, B, J, cell_centre_particles(nt, n, ppc), nt, n, stride, 2)
        no = [a * max(1, b // stride) for a, b in zip(nt, n)]
        assert out.shape == (11, no[2], no[1], no[0])
        for f, v in enumerate((1, 2, 3, 4, 5, 6)):
            assert np.all(out[f] == v)
        for f, v in enumerate((7, 8, 9)):
            assert np.all(out[6 + f] == v * stride**3)
        assert np.all(out[9] == ppc * stride**3) and np.all(out[10] == ppc * stride**3)


def test_spatially_varying_fields_and_placement():
    # test_spatially_varying_fields, test_multi_tile_placement: E_x = global x,
    # E_y = global y, B_z = global z, sampled at the stride points
    nt, n, stride = (2, 3, 2), (4, 6, 8), 2
    X = [a * b for a, b in zip(nt, n)]
    gx, gy, gz = np.meshgrid(*(np.arange(v, dtype=F32) for v in X), indexing="ij")
    E = [gx, gy, np.zeros(X, F32)]
    B = [np.zeros(X, F32), np.zeros(X, F32), gz]
    J = [np.ones(X, F32)] * 3
    parts = [[np.zeros((3, 0), F32)] for _ in range(12)]
    out = osnap.fields_pack(E, B, J, parts, nt, n, stride, 1)
    no = [a * (b // stride) for a, b in zip(nt, n)]
    assert out.shape == (10, no[2], no[1], no[0])
    ox = np.arange(no[0])
    xs = (ox // (n[0] // stride)) * n[0] + (ox % (n[0] // stride)) * stride
    assert np.array_equal(out[0][0, 0, :], xs.astype(F32))
    oy = np.arange(no[1])
    ys = (oy // (n[1] // stride)) * n[1] + (oy % (n[1] // stride)) * stride
    assert np.array_equal(out[1][0, :, 0], ys.astype(F32))
    oz = np.arange(no[2])
    zs = (oz // (n[2] // stride)) * n[2] + (oz % (n[2] // stride)) * stride
    assert np.array_equal(out[5][:, 0, 0], zs.astype(F32))
    assert np.all(out[6] == stride**3)


def test_j_volume_sum_order_and_values():
    # J summed over the stride^3 block: brute force against float64 sums
    rng = np.random.default_rng(3)
    nt, n, stride = (1, 1, 1), (6, 6, 6), 3
    J = [rng.standard_normal(n).astype(F32) for _ in range(3)]
    zero = [np.zeros(n, F32)] * 3
    out = osnap.fields_pack(zero, zero, J, [[]], nt, n, stride, 0)
    for c in range(3):
        ref = J[c].astype(np.float64).reshape(2, 3, 2, 3, 2, 3).sum(axis=(1, 3, 5)).transpose(2, 1, 0)
        assert np.allclose(out[6 + c], ref, rtol=1e-5, atol=1e-6)


def test_density_ignores_other_tiles_particles_outside_block():
    # a particle left outside its tile's block (before migration) is not counted
    nt, n = (1, 1, 1), (4, 4, 4)
    zero = [np.zeros(n, F32)] * 3
    p = np.array([[0.5, 3.9, -0.1, 4.2], [0.5, 0.5, 0.5, 0.5], [0.5, 0.5, 0.5, 0.5]], dtype=F32)
    out = osnap.fields_pack(zero, zero, zero, [[p]], nt, n, 1, 1)
    assert out[9].sum() == 2 and out[9][0, 0, 0] == 1 and out[9][0, 0, 3] == 1


def test_sample_plan():
    # test_all_particles_when_n_exceeds_total, test_sampling_reduces_count,
    # test_empty_species: per tile quota = round(count * n / total)
    assert osnap.sample_plan([512] * 4, 10 * 2048) == [(1, 512)] * 4
    plan = osnap.sample_plan([512, 512, 0, 100], 281)
    tot = sum(s for _, s in plan)
    assert plan[2] == (1, 0)
    assert 0 < tot <= 281 + 4 and abs(tot - 281) <= 281 * 0.5
    for (stride, k), c in zip(plan, [512, 512, 0, 100]):
        if k:
            assert k == c // stride and k <= c
    assert osnap.sample_plan([0, 0], 100) == [(1, 0), (1, 0)]
    assert osnap.sample_plan([3], 1) == [(3, 1)]


def spectra_of(vel, nt=(1, 1, 1), n=(8, 8, 8), nbins=50, umin=1e-2, umax=1e2, stride=1, nsp=1):
    parts = cell_centre_particles(nt, n, 1, nsp)
    px = [[p[0] for p in tp] for tp in parts]
    pv = [[np.tile(np.asarray(v, F32).reshape(3, 1), (1, tp[0].shape[1])) for v in vel[:nsp]] for tp in parts]
    return osnap.spectra_pack(px, pv, nt, n, stride, nbins, umin, umax, nsp)


def test_known_velocity_spectra():
    # test_known_velocity_u_spectrum: |u| = 5 in the bin of u = 5;
    # test_known_velocity_beta_spectrum: beta_x = 1/sqrt(2)
    out = spectra_of([(3.0, 4.0, 0.0)], nbins=100, umin=1e-2, umax=1e4)
    edges = np.logspace(np.log10(1e-2), np.log10(1e4), 101)
    tot = out[0].sum(axis=(0, 1, 2))
    assert tot[np.searchsorted(edges, 5.0) - 1] == 512 and tot.sum() == 512
    out = spectra_of([(1.0, 0.0, 0.0)], nbins=100)
    tot = out[1].sum(axis=(0, 1, 2))
    assert tot[np.searchsorted(np.linspace(-1, 1, 101), 1 / np.sqrt(2)) - 1] == 512
    # beta_y = beta_z = 0: the bin starting at 0
    assert out[2].sum(axis=(0, 1, 2))[50] == 512 and out[3].sum(axis=(0, 1, 2))[50] == 512


@pytest.mark.parametrize("vx, umin, umax, expected", [(100.0, 1e-2, 10.0, 49), (0.001, 1.0, 1e4, 0), (0.0, 1.0, 1e4, 0)])
def test_u_out_of_range_clamped(vx, umin, umax, expected):
    # test_u_out_of_range_clamped_to_boundary_bins (u = 0: log10 = -inf, bin 0)
    out = spectra_of([(vx, 0.0, 0.0)], nbins=50, umin=umin, umax=umax)
    tot = out[0].sum(axis=(0, 1, 2))
    assert tot[expected] == 512 and tot.sum() == 512


def test_multi_tile_x_resolved_and_species():
    # test_multi_tile_x_resolved, test_two_species_distinct_spectra: 2x1x1
    # tiles, stride 2: tile i fills x bins [4 i, 4 i + 4) only
    nt, n = (2, 1, 1), (8, 8, 8)
    parts = cell_centre_particles(nt, n, 1, 2)
    px = [[p[0] for p in tp] for tp in parts]
    pv = [[np.tile(np.array([[0.5 * (t + 1)], [0.0], [0.0]], F32), (1, tp[s].shape[1])) * (s + 1)
           for s in range(2)] for t, tp in enumerate(parts)]
    out = osnap.spectra_pack(px, pv, nt, n, 2, 40, 1e-2, 1e2, 2)
    assert out.shape == (8, 1, 1, 8, 40)
    for t in range(2):
        per_x = out[0, 0, 0].sum(axis=1)
        assert np.all(per_x[4 * t:4 * t + 4] == 512 / 4)
    u0 = out[0].sum(axis=(0, 1, 2))
    u1 = out[4].sum(axis=(0, 1, 2))
    assert u0.sum() == u1.sum() == 1024 and not np.array_equal(u0, u1)

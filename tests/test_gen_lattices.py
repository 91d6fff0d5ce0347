"""The generator's triangular, honeycomb and kagome lattices.

The reference's own checks (test/test_tight_binding.py:7-174: shape, dtype,
real at zero flux, tp refused) are restated, and pinned further by what the
physics fixes: the zero-flux spectrum equals the closed-form band structure
of each lattice, every site has the lattice's coordination number, and with
flux every elementary loop carries the same phase, the fraction of the cell's
flux nflux/(Nx Ny) that its area holds (triangles 1/2, kagome triangles 1/8,
hexagons 1).  The hs_init / hs_end trajectories of test/hs_ref.py pin the
zero-flux hopping matrices of all four lattices (test_oracle_golden.py).
"""
import numpy as np
import pytest

from dqmc_amd import gen

Nx, Ny = 6, 4


def _bands(geometry):
    """Closed-form tight-binding bands, t = 1, on the Nx x Ny k-grid."""
    k1 = 2 * np.pi * np.arange(Nx)[:, None] / Nx
    k2 = 2 * np.pi * np.arange(Ny)[None, :] / Ny
    if geometry == "triangular":
        # neighbours +-a1, +-a2, +-(a1 - a2)
        return (-2 * (np.cos(k1) + np.cos(k2) + np.cos(k1 - k2))).ravel()
    if geometry == "honeycomb":
        # A hops to B of its cell and of the cells +a1, +a2
        f = np.abs(1 + np.exp(1j * k1) + np.exp(1j * k2)).ravel()
        return np.concatenate([f, -f])
    if geometry == "kagome":
        # flat band at +2t; -t (1 +- sqrt(4 sum cos^2(k.a/2) - 3)), a3 = a1 - a2
        c = np.cos(k1 / 2) ** 2 + np.cos(k2 / 2) ** 2 + np.cos((k1 - k2) / 2) ** 2
        r = np.sqrt(np.maximum(4 * c - 3, 0)).ravel()
        return np.concatenate([np.full(r.size, 2.0), -(1 + r), -(1 - r)])
    raise ValueError(geometry)


HOP = {"triangular": gen.hopping_triangular, "honeycomb": gen.hopping_honeycomb,
       "kagome": gen.hopping_kagome}
COORD = {"triangular": 6, "honeycomb": 3, "kagome": 4}


@pytest.mark.parametrize("geometry", sorted(HOP))
def test_zero_flux_matrix_is_real_with_textbook_spectrum(geometry):
    K, P = HOP[geometry](Nx, Ny, nflux=0)
    N = gen.NORB[geometry] * Nx * Ny
    assert K.shape == (N, N) and K.dtype == np.complex128
    assert np.abs(P.imag).max() < 1e-10 and np.abs(K.imag).max() < 1e-10
    assert np.allclose(K, K.T.conj())
    ev = np.sort(np.linalg.eigvalsh(K))
    assert np.abs(ev - np.sort(_bands(geometry))).max() < 1e-12
    deg = (np.abs(K) > 1e-12).sum(axis=0)
    assert np.all(deg == COORD[geometry])


def _loops(K, length):
    """Simple closed loops of the given length on the bond graph of K, one per
    vertex set, in the traversal order first found."""
    n = K.shape[0]
    nb = [np.nonzero(np.abs(K[:, i]) > 1e-12)[0] for i in range(n)]
    seen = {}
    for s in range(n):
        stack = [[s]]
        while stack:
            p = stack.pop()
            if len(p) == length:
                if s in nb[p[-1]] and frozenset(p) not in seen:
                    seen[frozenset(p)] = p
                continue
            stack.extend(p + [w] for w in nb[p[-1]] if w not in p and w > s)
    return list(seen.values())


def _phase(K, loop):
    z = 1.0
    for a, b in zip(loop, loop[1:] + loop[:1]):
        z *= -K[b, a]  # hop a -> b, t = 1
    assert abs(abs(z) - 1) < 1e-12
    return np.angle(z) / (2 * np.pi)


@pytest.mark.parametrize("nflux", [1, 3])
def test_uniform_field_triangular(nflux):
    K, _ = gen.hopping_triangular(Nx, Ny, nflux=nflux)
    cell = lambda x, y: (x % Nx) + Nx * (y % Ny)
    ph = []
    for y in range(Ny):
        for x in range(Nx):
            # both triangles of the cell, counterclockwise (a1 = (1, 0), a2 at 60 degrees)
            ph.append(_phase(K, [cell(x, y), cell(x + 1, y), cell(x, y + 1)]))
            ph.append(_phase(K, [cell(x + 1, y), cell(x + 1, y + 1), cell(x, y + 1)]))
    assert np.allclose(ph, nflux / (2 * Nx * Ny), atol=1e-12)


@pytest.mark.parametrize("nflux", [1, 3])
def test_uniform_field_kagome(nflux):
    K, _ = gen.hopping_kagome(Nx, Ny, nflux=nflux)
    tris = _loops(K, 3)
    assert len(tris) == 2 * Nx * Ny
    # traversed A -> B -> C (ascending site index): the same sense for both
    # triangles of the cell
    ph = [_phase(K, sorted(t)) for t in tris]
    assert np.allclose(np.abs(ph), nflux / (8 * Nx * Ny), atol=1e-12)
    assert np.ptp(ph) < 1e-12


@pytest.mark.parametrize("nflux", [1, 3])
def test_uniform_field_honeycomb(nflux):
    K, _ = gen.hopping_honeycomb(Nx, Ny, nflux=nflux)
    hexes = _loops(K, 6)
    assert len(hexes) == Nx * Ny
    ph = np.array([_phase(K, h) for h in hexes])
    # loop sense is not fixed by the traversal: compare magnitudes
    assert np.allclose(np.abs(ph), nflux / (Nx * Ny), atol=1e-12)


@pytest.mark.parametrize("geometry", sorted(HOP))
def test_maps_bin_pairs_by_displacement_and_orbitals(geometry):
    s = gen.create_batch(seed=7, geometry=geometry, Nx=Nx, Ny=Ny)[0]
    p = s["params"]
    Norb, Nc = gen.NORB[geometry], Nx * Ny
    N = Norb * Nc
    assert int(p["N"]) == N
    site = np.arange(N)
    x, y, o = site % Nx, (site // Nx) % Ny, site // Nc
    key = {}
    for j in range(N):
        for i in range(N):
            cls = ((x[i] - x[j]) % Nx, (y[i] - y[j]) % Ny, o[i], o[j])
            assert key.setdefault(int(p["map_ij"][j, i]), cls) == cls
    assert len(key) == int(p["num_ij"]) == Norb * Norb * Nc
    assert np.all(p["degen_ij"] == Nc)
    assert np.all(p["degen_i"] == Nc) and int(p["num_i"]) == Norb
    assert np.array_equal(p["map_i"], o)


@pytest.mark.parametrize("geometry", sorted(HOP))
def test_options_without_a_definition_are_refused(geometry):
    with pytest.raises(Exception):
        gen.create_1(geometry=geometry, Nx=4, Ny=4, tp=0.2)

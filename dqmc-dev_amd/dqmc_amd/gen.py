"""Simulation-input generator for the single-band Hubbard DQMC engine.

Builds every array of the ``.h5`` simulation layout (``/params``, ``/state``,
``/meas_eqlt``, ``/meas_uneqlt``) as numpy arrays, with the names, dtypes and
C-order shapes that util/gen_1band_unified_hub.py writes, so a file made from
this dict is read by the engine exactly like one made by the reference
generator.  Covered: square lattice with periodic boundaries (bc=1), t, t', mu,
h, U, dt, L, Peierls flux (nflux, gauge alpha=1/2) and twists, every
measurement map the square geometry defines, and the seeded RNG / initial HS
field.  The triangular hopping matrix is provided for trajectory tests only
(no triangular measurement maps).

Reference semantics followed (file:line of katherineding/dqmc-dev):
- RNG seeding / draws / jump: util/gen_util_shared.py:11-66
- initial HS field (rand_uint >> 63 over (L, N)): util/gen_1band_unified_hub.py:325-328
- maps and degeneracies: util/gen_1band_unified_hub.py:40-263, 330-633
- hopping + Peierls phases: util/tight_binding.py:4-136, 182-224
- two-hop (thermal) phases: util/gen_1band_unified_hub.py:662-691
- K_up/K_dn, exponentials, lambda, del: util/gen_1band_unified_hub.py:903-927
- dataset names/dtypes: util/gen_1band_unified_hub.py:929-1160
- batches with rand_jump per file: util/gen_1band_unified_hub.py:1163-1195
"""
import math

import numpy as np
from scipy.linalg import expm

MASK64 = (1 << 64) - 1
XS_MULT = 1181783497276652981
# xorshift1024* jump polynomial (2^512 draws), Vigna's published constants
JUMP = (
    0x84242F96ECA9C41D, 0xA3C65B8776F96855, 0x5B34A39F070B5837, 0x4489AFFCE4F31A1E,
    0x2FFEEB0A48316F40, 0xDC2D9891FE68C022, 0x3659132BB12FEA70, 0xAAC17D8EFA43CAB8,
    0xC4CB815590989B13, 0x5EE975283D71C93B, 0x691548C86C1BD540, 0x7910C41D10A1E6A5,
    0x0B5FC64563B3E2A8, 0x047F7684E9FC949D, 0xB99181F2D8F685CA, 0x284600E3F30E38C3,
)


# ----------------------------------------------------------------------------
# RNG: splitmix64 seeding, xorshift1024* draws, jump (util/gen_util_shared.py)
# ----------------------------------------------------------------------------
def splitmix64_seed(x):
    """17-word xorshift1024* state from splitmix64(x); word 16 is the index p."""
    x &= MASK64
    s = []
    for _ in range(16):
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        s.append(z ^ (z >> 31))
    return s + [0]


class XorShift1024Star:
    """Pure-Python xorshift1024* over a 17-word state (list of ints)."""

    def __init__(self, state):
        self.s = [int(v) & MASK64 for v in state]

    def next(self):
        s = self.s
        s0 = s[s[16]]
        p = (s[16] + 1) & 15
        s[16] = p
        s1 = s[p]
        s1 ^= (s1 << 31) & MASK64
        s[p] = s1 ^ s0 ^ (s1 >> 11) ^ (s0 >> 30)
        return (s[p] * XS_MULT) & MASK64

    def jump(self):
        s = self.s
        t = [0] * 16
        for word in JUMP:
            for b in range(64):
                if (word >> b) & 1:
                    p = s[16]
                    for j in range(16):
                        t[j] ^= s[(j + p) & 15]
                self.next()
        p = s[16]
        for j in range(16):
            s[(j + p) & 15] = t[j]

    def state(self):
        return np.array(self.s, dtype=np.uint64)


def init_hs_field(rng, L, N):
    hs = np.empty((L, N), dtype=np.int32)
    for l in range(L):
        for i in range(N):
            hs[l, i] = rng.next() >> 63
    return hs


# ----------------------------------------------------------------------------
# Hopping matrices (util/tight_binding.py)
# ----------------------------------------------------------------------------
def peierls_matrix(Nx, Ny, shape, nflux, alpha=0.5, jr_orbit=(0.0, 0.0), ir_orbit=(0.0, 0.0)):
    """Peierls phase matrix exp(i*2*pi*nflux/(A*Nx*Ny) * phi); phi[j, i] for the
    hop i -> j with the symmetric-gauge wrap correction (tight_binding.py:4-78).
    ir_orbit / jr_orbit: positions of the start / end orbital in the unit cell
    (the honeycomb and kagome sublattices)."""
    beta = 1.0 - alpha
    if shape == "triangular":
        a1 = (0.5, math.sqrt(3) / 2)
        a2 = (-0.5, math.sqrt(3) / 2)
        unit_area = math.sqrt(3) / 2
    else:
        a1 = (1.0, 0.0)
        a2 = (0.0, 1.0)
        unit_area = 1.0
    prefactor = 2 * np.pi * nflux / (unit_area * Nx * Ny)
    N = Nx * Ny
    phi = np.zeros((N, N))
    for dy in range((1 - Ny) // 2, (1 + Ny) // 2):
        for dx in range((1 - Nx) // 2, (1 + Nx) // 2):
            for iy in range(Ny):
                for ix in range(Nx):
                    jy = iy + dy
                    jjy = jy % Ny
                    jx = ix + dx
                    jjx = jx % Nx
                    o1 = jx - jjx
                    o2 = jy - jjy
                    ir = (ix * a1[0] + iy * a2[0] + ir_orbit[0], ix * a1[1] + iy * a2[1] + ir_orbit[1])
                    jr = (jx * a1[0] + jy * a2[0] + jr_orbit[0], jx * a1[1] + jy * a2[1] + jr_orbit[1])
                    jjr = (jjx * a1[0] + jjy * a2[0] + jr_orbit[0],
                           jjx * a1[1] + jjy * a2[1] + jr_orbit[1])
                    dr = (jr[0] - ir[0], jr[1] - ir[1])
                    mr = ((ir[0] + jr[0]) / 2, (ir[1] + jr[1]) / 2)
                    A_off = (-alpha * (o1 * a1[1] + o2 * a2[1]),
                             beta * (o1 * a1[0] + o2 * a2[0]))
                    interior = -alpha * mr[1] * dr[0] + beta * mr[0] * dr[1]
                    wrap = A_off[0] * (0.5 * (jr[0] + jjr[0])) + A_off[1] * (0.5 * (jr[1] + jjr[1]))
                    extra = o1 * o2 * (beta * a2[0] * a1[1] - alpha * a1[0] * a2[1])
                    phi[jjx + Nx * jjy, ix + Nx * iy] = interior + (-1) * (wrap + extra)
    return np.exp(1j * prefactor * phi)


def hopping_square(Nx, Ny, t=1.0, tp=0.0, tpp=0.0, nflux=0, alpha=0.5, twistx=0.0, twisty=0.0):
    """Periodic square lattice hopping -t_ij exp(i phi_ij) (tight_binding.py:81-136)."""
    N = Nx * Ny
    tij = np.zeros((N, N), dtype=np.complex128)
    ex, ey = np.exp(1j * twistx), np.exp(1j * twisty)
    for iy in range(Ny):
        for ix in range(Nx):
            iy1 = (iy + 1) % Ny
            ix1 = (ix + 1) % Nx
            tij[ix1 + Nx * iy, ix + Nx * iy] += -t * ex
            tij[ix + Nx * iy, ix1 + Nx * iy] += -t * np.exp(-1j * twistx)
            tij[ix + Nx * iy1, ix + Nx * iy] += -t * ey
            tij[ix + Nx * iy, ix + Nx * iy1] += -t * np.exp(-1j * twisty)
            tij[ix + Nx * iy, ix1 + Nx * iy1] += -tp * np.exp(-1j * twistx - 1j * twisty)
            tij[ix1 + Nx * iy1, ix + Nx * iy] += -tp * np.exp(+1j * twistx + 1j * twisty)
            tij[ix1 + Nx * iy, ix + Nx * iy1] += -tp * np.exp(+1j * twistx - 1j * twisty)
            tij[ix + Nx * iy1, ix1 + Nx * iy] += -tp * np.exp(-1j * twistx + 1j * twisty)
            iy2 = (iy + 2) % Ny
            ix2 = (ix + 2) % Nx
            tij[ix2 + Nx * iy, ix + Nx * iy] += -tpp * np.exp(1j * twistx * 2)
            tij[ix + Nx * iy, ix2 + Nx * iy] += -tpp * np.exp(-1j * twistx * 2)
            tij[ix + Nx * iy2, ix + Nx * iy] += -tpp * np.exp(1j * twisty * 2)
            tij[ix + Nx * iy, ix + Nx * iy2] += -tpp * np.exp(-1j * twisty * 2)
    peierls = peierls_matrix(Nx, Ny, "square", nflux, alpha)
    K = tij * peierls
    if nflux > 0 or twistx != 0 or twisty != 0:
        assert np.linalg.norm(K - K.T.conj()) < 1e-10
    return K, peierls


def hopping_triangular(Nx, Ny, t=1.0, nflux=0, alpha=0.5):
    """Periodic triangular lattice hopping (tight_binding.py:182-224)."""
    N = Nx * Ny
    kij = np.zeros((N, N), dtype=np.complex128)
    for iy in range(Ny):
        for ix in range(Nx):
            iy1 = (iy + 1) % Ny
            ix1 = (ix + 1) % Nx
            kij[ix1 + Nx * iy, ix + Nx * iy] += -t
            kij[ix + Nx * iy, ix1 + Nx * iy] += -t
            kij[ix + Nx * iy1, ix + Nx * iy] += -t
            kij[ix + Nx * iy, ix + Nx * iy1] += -t
            kij[ix1 + Nx * iy, ix + Nx * iy1] += -t
            kij[ix + Nx * iy1, ix1 + Nx * iy] += -t
    peierls = peierls_matrix(Nx, Ny, "triangular", nflux, alpha)
    return kij * peierls, peierls


def hopping_honeycomb(Nx, Ny, t=1.0, nflux=0, alpha=0.5):
    """Periodic honeycomb lattice, orbitals A (0) and B (1) per triangular-lattice
    cell, site index x + Nx y + Nx Ny o (tight_binding.py:227-324).  Each A
    site hops to the B sites of its own cell and of the cells one step along
    either primitive vector.  Returns (K, P): the A-A and B-B blocks of P are
    the triangular-lattice phases, the B<-A block the phases of hops to a site
    displaced by the B offset (0, -1/sqrt 3), the A<-B block its conjugate
    transpose; the off-diagonal blocks of K carry them."""
    Nc = Nx * Ny
    N = 2 * Nc
    tij = np.zeros((N, N), dtype=np.complex128)
    for iy in range(Ny):
        for ix in range(Nx):
            a = ix + Nx * iy
            for bx, by in ((ix, (iy + 1) % Ny), ((ix + 1) % Nx, iy), (ix, iy)):
                b = bx + Nx * by + Nc
                tij[b, a] += -t
                tij[a, b] += -t
    pd = peierls_matrix(Nx, Ny, "triangular", nflux, alpha)
    po = peierls_matrix(Nx, Ny, "triangular", nflux, alpha, jr_orbit=(0.0, -1.0 / math.sqrt(3)))
    P = np.empty((N, N), dtype=np.complex128)
    P[:Nc, :Nc] = pd
    P[Nc:, Nc:] = pd
    P[:Nc, Nc:] = po.T.conj()
    P[Nc:, :Nc] = po
    K = tij.copy()
    K[:Nc, :Nc] *= pd
    K[Nc:, Nc:] *= pd
    K[:Nc, Nc:] *= po.T.conj()
    K[Nc:, :Nc] *= po
    if nflux != 0:
        assert np.linalg.norm(K - K.T.conj()) < 1e-10
    return K, P


def hopping_kagome(Nx, Ny, t=1.0, nflux=0, alpha=0.5):
    """Periodic kagome lattice, orbitals A (0), B (1), C (2) per triangular-
    lattice cell, site index x + Nx y + Nx Ny o (tight_binding.py:330-452).
    Bonds: A-B within the cell and A(x, y) - B(x, y+1); A-C within the cell and
    C(x, y) - A(x-1, y); B-C within the cell and B(x, y) - C(x+1, y-1).  The
    sublattice offsets B (1/4, -sqrt3/4) and C (-1/4, -sqrt3/4) enter the
    phases of the inter-orbital blocks as in the reference."""
    Nc = Nx * Ny
    N = 3 * Nc
    tij = np.zeros((N, N), dtype=np.complex128)

    def bond(i, j):
        tij[i, j] += -t
        tij[j, i] += -t

    for iy in range(Ny):
        for ix in range(Nx):
            c = ix + Nx * iy
            bond(ix + Nx * ((iy + 1) % Ny) + Nc, c)                    # B(x, y+1) - A(x, y)
            bond(c + Nc, c)                                             # B - A, same cell
            bond(((ix - 1) % Nx) + Nx * iy, c + 2 * Nc)                 # A(x-1, y) - C(x, y)
            bond(c, c + 2 * Nc)                                         # A - C, same cell
            bond(((ix + 1) % Nx) + Nx * ((iy - 1) % Ny) + 2 * Nc, c + Nc)  # C(x+1, y-1) - B
            bond(c + 2 * Nc, c + Nc)                                    # C - B, same cell
    d4 = (0.25, -math.sqrt(3) / 4)   # B relative to A
    d3 = (-0.25, -math.sqrt(3) / 4)  # C relative to A
    p01 = peierls_matrix(Nx, Ny, "triangular", nflux, alpha, jr_orbit=d4)
    p20 = peierls_matrix(Nx, Ny, "triangular", nflux, alpha, ir_orbit=d3)
    p12 = peierls_matrix(Nx, Ny, "triangular", nflux, alpha, jr_orbit=d3, ir_orbit=d4)
    pd = peierls_matrix(Nx, Ny, "triangular", nflux, alpha)
    A, B, C = slice(0, Nc), slice(Nc, 2 * Nc), slice(2 * Nc, 3 * Nc)
    P = np.empty((N, N), dtype=np.complex128)
    for o in (A, B, C):
        P[o, o] = pd
    P[A, B], P[B, A] = p01.T.conj(), p01
    P[C, A], P[A, C] = p20.T.conj(), p20
    P[C, B], P[B, C] = p12, p12.T.conj()
    K = tij.copy()
    K[B, A] *= p01
    K[A, B] *= p01.T.conj()
    K[C, A] *= p20.T.conj()
    K[A, C] *= p20
    K[C, B] *= p12
    K[B, C] *= p12.T.conj()
    if nflux != 0:
        assert np.linalg.norm(K - K.T.conj()) < 1e-10
    return K, P


NORB = {"square": 1, "triangular": 1, "honeycomb": 2, "kagome": 3}


def _orbital_maps(geometry, Nx, Ny, trans_sym):
    """Maps of the triangular, honeycomb and kagome lattices
    (gen_1band_unified_hub.py:40-263, 697-903): sites by orbital, site pairs by
    (displacement of the cells, orbital of each site), plaquettes (the square
    lattice's two triangles per cell on the triangular lattice; the A-C-B
    triangles on kagome; none on honeycomb).  The reference defines no bonds,
    two-hop bonds or bond maps for these lattices (placeholders of zeros); the
    same placeholders are written here, and the measurements that need them
    are refused by create_1."""
    Norb = NORB[geometry]
    Nc = Nx * Ny
    N = Norb * Nc
    x = np.arange(N) % Nx
    y = (np.arange(N) // Nx) % Ny
    o = np.arange(N) // Nc
    m = {}
    if trans_sym:
        map_i = o.astype(np.int32)
        degen_i = np.full(Norb, Nc, dtype=np.int32)
        # bin of the pair (j, i): cell displacement i - j and both orbitals
        k = ((x[None, :] - x[:, None]) % Nx + Nx * ((y[None, :] - y[:, None]) % Ny)
             + Nc * o[None, :] + N * o[:, None])
        num_ij = Norb * Norb * Nc
    else:
        map_i = np.arange(N, dtype=np.int32)
        degen_i = np.ones(N, dtype=np.int32)
        k = np.arange(N)[None, :] + N * np.arange(N)[:, None]
        num_ij = N * N
    map_ij = k.astype(np.int32)  # map_ij[j, i]
    degen_ij = np.bincount(map_ij.reshape(-1), minlength=num_ij).astype(np.int32)
    assert np.all(degen_ij == degen_ij[0])
    m.update(map_i=map_i, degen_i=degen_i, num_i=int(map_i.max()) + 1, map_ij=map_ij,
             degen_ij=degen_ij, num_ij=num_ij)
    ppc = {"triangular": 2, "honeycomb": 1, "kagome": 2}[geometry]
    num_plaq = ppc * Nc
    if trans_sym:
        map_plaq = np.repeat(np.arange(ppc, dtype=np.int32), Nc)
        degen_plaq = np.full(ppc, Nc, dtype=np.int32)
    else:
        map_plaq = np.arange(num_plaq, dtype=np.int32)
        degen_plaq = np.ones(num_plaq, dtype=np.int32)
    plaqs = np.zeros((3, num_plaq), dtype=np.int32)
    c = np.arange(Nc)
    cx, cy = c % Nx, c // Nx
    cell = lambda a, b: (a % Nx) + Nx * (b % Ny)
    if geometry == "triangular":
        plaqs[:, :Nc] = (c, cell(cx + 1, cy), cell(cx, cy + 1))
        plaqs[:, Nc:] = (cell(cx + 1, cy), cell(cx + 1, cy + 1), cell(cx, cy + 1))
    elif geometry == "kagome":
        plaqs[:, :Nc] = (c, c + 2 * Nc, c + Nc)
        plaqs[:, Nc:] = (c, cell(cx + 1, cy) + 2 * Nc, cell(cx, cy + 1) + Nc)
    m.update(plaqs=plaqs, map_plaq=map_plaq, degen_plaq=degen_plaq, num_plaq=num_plaq,
             num_plaq_accum=int(map_plaq.max()) + 1, plaq_per_cell=ppc)
    bps = {"triangular": 3, "honeycomb": 3, "kagome": 4}[geometry]
    b2ps = 4
    num_b, num_b2 = bps * N, b2ps * N
    for name, per, n in (("b", bps, num_b), ("b2", b2ps, num_b2)):
        if trans_sym:
            mp = np.repeat(np.arange(per, dtype=np.int32), N)
            dg = np.full(per, N, dtype=np.int32)
        else:
            mp = np.arange(n, dtype=np.int32)
            dg = np.ones(n, dtype=np.int32)
        m.update({"map_" + name: mp, "degen_" + name: dg, "num_%s_accum" % name: int(mp.max()) + 1})
    cnt = {"bs": (bps * N if trans_sym else num_b * N, (N, num_b)),
           "bb": (bps * bps * N if trans_sym else num_b * num_b, (num_b, num_b)),
           "b2b2": (b2ps * b2ps * N if trans_sym else num_b2 * num_b2, (num_b2, num_b2)),
           "bb2": (bps * b2ps * N if trans_sym else num_b * num_b2, (num_b, num_b2)),
           "b2b": (b2ps * bps * N if trans_sym else num_b2 * num_b, (num_b2, num_b))}
    for name, (n, shape) in cnt.items():
        m["num_" + name] = n
        m["map_" + name] = np.zeros(shape, dtype=np.int32)
        m["degen_" + name] = np.zeros(max(n, 1), dtype=np.int32)
    m.update(bps=bps, b2ps=b2ps, num_b=num_b, num_b2=num_b2,
             bonds=np.zeros((2, num_b), dtype=np.int32),
             bond2s=np.zeros((2, num_b2), dtype=np.int32))
    return m


# ----------------------------------------------------------------------------
# Measurement maps for the square lattice (gen_1band_unified_hub.py:40-633)
# ----------------------------------------------------------------------------
def _square_maps(Nx, Ny, tp, trans_sym):
    N = Nx * Ny
    ix = np.arange(N) % Nx
    iy = np.arange(N) // Nx
    site = lambda x, y: (x % Nx) + Nx * (y % Ny)

    m = {}
    # 1-site map
    if trans_sym:
        map_i = np.zeros(N, dtype=np.int32)
        degen_i = np.full(1, N, dtype=np.int32)
    else:
        map_i = np.arange(N, dtype=np.int32)
        degen_i = np.ones(N, dtype=np.int32)
    m["map_i"], m["degen_i"], m["num_i"] = map_i, degen_i, int(map_i.max()) + 1

    # plaquettes (two triangles per cell)
    num_plaq = 2 * N
    if trans_sym:
        map_plaq = np.repeat(np.arange(2, dtype=np.int32), N)
        degen_plaq = np.full(2, N, dtype=np.int32)
    else:
        map_plaq = np.arange(num_plaq, dtype=np.int32)
        degen_plaq = np.ones(num_plaq, dtype=np.int32)
    plaqs = np.zeros((3, num_plaq), dtype=np.int32)
    plaqs[0, :N] = np.arange(N)
    plaqs[1, :N] = site(ix + 1, iy)
    plaqs[2, :N] = site(ix, iy + 1)
    plaqs[0, N:] = site(ix + 1, iy)
    plaqs[1, N:] = site(ix + 1, iy + 1)
    plaqs[2, N:] = site(ix, iy + 1)
    m.update(plaqs=plaqs, map_plaq=map_plaq, degen_plaq=degen_plaq, num_plaq=num_plaq,
             num_plaq_accum=int(map_plaq.max()) + 1, plaq_per_cell=2)

    # 1-bonds
    bps = 4 if tp != 0.0 else 2
    num_b = bps * N
    if trans_sym:
        map_b = np.repeat(np.arange(bps, dtype=np.int32), N)
        degen_b = np.full(bps, N, dtype=np.int32)
    else:
        map_b = np.arange(num_b, dtype=np.int32)
        degen_b = np.ones(num_b, dtype=np.int32)
    bonds = np.zeros((2, num_b), dtype=np.int32)
    bonds[0, :N] = np.arange(N)
    bonds[1, :N] = site(ix + 1, iy)
    bonds[0, N:2 * N] = np.arange(N)
    bonds[1, N:2 * N] = site(ix, iy + 1)
    if bps == 4:
        bonds[0, 2 * N:3 * N] = np.arange(N)
        bonds[1, 2 * N:3 * N] = site(ix + 1, iy + 1)
        bonds[0, 3 * N:] = site(ix + 1, iy)
        bonds[1, 3 * N:] = site(ix, iy + 1)
    m.update(bps=bps, num_b=num_b, map_b=map_b, degen_b=degen_b, bonds=bonds,
             num_b_accum=int(map_b.max()) + 1)

    # 2-bonds (pairs of sites two hops apart)
    b2ps = 12 if tp != 0.0 else 4
    num_b2 = b2ps * N
    if trans_sym:
        map_b2 = np.repeat(np.arange(b2ps, dtype=np.int32), N)
        degen_b2 = np.full(b2ps, N, dtype=np.int32)
    else:
        map_b2 = np.arange(num_b2, dtype=np.int32)
        degen_b2 = np.ones(num_b2, dtype=np.int32)
    b2def = [((0, 0), (2, 0)), ((0, 0), (0, 2)), ((0, 0), (1, 1)), ((1, 0), (0, 1))]
    if b2ps == 12:
        b2def += [((0, 0), (2, 1)), ((0, 0), (1, 2)), ((2, 0), (0, 1)), ((1, 0), (0, 2)),
                  ((0, 0), (0, 1)), ((0, 0), (1, 0)), ((0, 0), (2, 2)), ((2, 0), (0, 2))]
    bond2s = np.zeros((2, num_b2), dtype=np.int32)
    for t_, ((x0, y0), (x1, y1)) in enumerate(b2def):
        bond2s[0, t_ * N:(t_ + 1) * N] = site(ix + x0, iy + y0)
        bond2s[1, t_ * N:(t_ + 1) * N] = site(ix + x1, iy + y1)
    m.update(b2ps=b2ps, num_b2=num_b2, map_b2=map_b2, degen_b2=degen_b2, bond2s=bond2s,
             num_b2_accum=int(map_b2.max()) + 1)

    # 2-site map: map_ij[j, i] = (i - j) translation class
    jj, ii = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    if trans_sym:
        num_ij = N
        k = ((ii % Nx - jj % Nx) % Nx) + Nx * ((ii // Nx - jj // Nx) % Ny)
    else:
        num_ij = N * N
        k = ii + N * jj
    map_ij = k.astype(np.int32)
    m.update(map_ij=map_ij, num_ij=num_ij,
             degen_ij=np.bincount(map_ij.ravel(), minlength=num_ij).astype(np.int32))

    def pair_map(nb_j, nb_i, stride):
        # out[j + N*jb, i + N*ib] = map_ij[j, i] + num_ij * (ib + stride*jb)
        jb = np.arange(nb_j)[:, None, None, None]
        ib = np.arange(nb_i)[None, None, :, None]
        out = map_ij[None, :, None, :] + num_ij * (ib + stride * jb)
        return out.reshape(nb_j * N, nb_i * N).astype(np.int32)

    if trans_sym:
        map_bs = (map_ij[:, None, :] + num_ij * np.arange(bps)[None, :, None]).reshape(N, bps * N)
        map_bs = map_bs.astype(np.int32)
        map_bb = pair_map(bps, bps, bps)
        map_b2b2 = pair_map(b2ps, b2ps, b2ps)
        map_bb2 = pair_map(bps, b2ps, b2ps)
        map_b2b = pair_map(b2ps, bps, bps)
    else:  # the reference only supports trans_sym maps for the bond sets on the square lattice
        map_bs = (map_ij[:, None, :] + num_ij * np.arange(bps)[None, :, None]).reshape(N, bps * N)
        map_bs = map_bs.astype(np.int32)
        map_bb = pair_map(bps, bps, bps)
        map_b2b2 = pair_map(b2ps, b2ps, b2ps)
        map_bb2 = pair_map(bps, b2ps, b2ps)
        map_b2b = pair_map(b2ps, bps, bps)
    for name, arr in (("bs", map_bs), ("bb", map_bb), ("b2b2", map_b2b2), ("bb2", map_bb2),
                      ("b2b", map_b2b)):
        num = int(arr.max()) + 1
        m["map_" + name] = arr
        m["num_" + name] = num
        m["degen_" + name] = np.bincount(arr.ravel(), minlength=num).astype(np.int32)

    # two-hop paths (start, middle, end) per bond2 type, for the thermal phases
    hop2 = [((0, 0), (1, 0), (2, 0)), ((0, 0), (0, 1), (0, 2)), ((0, 0), (1, 0), (1, 1)),
            ((0, 0), (0, 1), (1, 1)), ((1, 0), (1, 1), (0, 1)), ((1, 0), (0, 0), (0, 1))]
    if b2ps == 12:
        hop2 += [((0, 0), (1, 0), (2, 1)), ((0, 0), (1, 1), (2, 1)), ((0, 0), (1, 1), (1, 2)),
                 ((0, 0), (0, 1), (1, 2)), ((2, 0), (1, 1), (0, 1)), ((2, 0), (1, 0), (0, 1)),
                 ((1, 0), (0, 1), (0, 2)), ((1, 0), (1, 1), (0, 2)), ((0, 0), (1, 1), (0, 1)),
                 ((0, 0), (1, 0), (0, 1)), ((0, 0), (-1, 1), (0, 1)), ((0, 0), (-1, 0), (0, 1)),
                 ((0, 0), (1, 1), (1, 0)), ((0, 0), (0, 1), (1, 0)), ((0, 0), (1, -1), (1, 0)),
                 ((0, 0), (0, -1), (1, 0)), ((0, 0), (1, 1), (2, 0)), ((0, 0), (1, -1), (2, 0)),
                 ((0, 0), (1, 1), (0, 2)), ((0, 0), (-1, 1), (0, 2)), ((0, 0), (1, 1), (2, 2)),
                 ((2, 0), (1, 1), (0, 2))]
        bond_hop = {0: [0, 22, 23], 1: [1, 24, 25], 8: [14, 15, 16, 17], 9: [18, 19, 20, 21],
                    10: [26], 11: [27]}
        for b in range(2, 8):
            bond_hop[b] = [2 * b - 2, 2 * b - 1]
    else:
        bond_hop = {0: [0], 1: [1], 2: [2, 3], 3: [4, 5]}
    m["hop2_mid"] = [np.asarray(site(ix + mid[0], iy + mid[1])) for (_, mid, _) in hop2]
    m["bond_hop"] = bond_hop
    return m


def _thermal_phases(m, peierls, tp, N):
    b2ps = m["b2ps"]
    bond2s = m["bond2s"]
    out = np.ones((b2ps, N), dtype=np.complex128)
    for i in range(N):
        for bt in range(b2ps):
            i0 = bond2s[0, i + bt * N]
            i2 = bond2s[1, i + bt * N]
            types = m["bond_hop"][bt]
            pp = 0
            if b2ps == 12 and bt in (0, 1):
                i1 = m["hop2_mid"][types[0]][i]
                pp += peierls[i0, i1] * peierls[i1, i2]
                for ht in types[1:]:
                    i1 = m["hop2_mid"][ht][i]
                    pp += tp * tp * peierls[i0, i1] * peierls[i1, i2]
            else:
                for ht in types:
                    i1 = m["hop2_mid"][ht][i]
                    pp += peierls[i0, i1] * peierls[i1, i2]
            out[bt, i] = pp
    return out


# ----------------------------------------------------------------------------
# create_1 / create_batch
# ----------------------------------------------------------------------------
MEAS_TOGGLES = ("meas_bond_corr", "meas_thermal", "meas_2bond_corr", "meas_energy_corr",
                "meas_nematic_corr", "meas_chiral", "meas_local_JQ", "meas_pair_bb_only",
                "meas_gen_suscept")


def create_1(init_rng=None, geometry="square", bc=1, Nx=4, Ny=4, mu=0.0, tp=0.0, tpp=0.0,
             U=6.0, dt=0.1, L=40, nflux=0, h=0.0, n_delay=16, n_matmul=8,
             n_sweep_warm=200, n_sweep_meas=2000, period_eqlt=8, period_uneqlt=0,
             trans_sym=1, checkpoint_every=10000, twistx=0.0, twisty=0.0, alpha=0.5,
             **toggles):
    """Return the simulation contents as {"params": {...}, "state": {...},
    "meas_eqlt": {...}, "meas_uneqlt": {...}, "metadata": {...}}.

    Defaults and assertions match util/gen_1band_unified_hub.py:265-310.
    alpha is the gauge of the Peierls phases: 1/2 (symmetric), the value the
    reference's generator always passes (util/gen_1band_unified_hub.py:643),
    or 1 (Landau).  With t' = 0 the two are related by a site-diagonal unitary
    and give the same gauge-invariant physics."""
    for k in toggles:
        if k not in MEAS_TOGGLES:
            raise TypeError("unknown argument %r" % k)
    tog = {k: int(toggles.get(k, 0)) for k in MEAS_TOGGLES}
    assert L % n_matmul == 0 and L % period_eqlt == 0
    if bc != 1:
        raise NotImplementedError("only periodic boundaries (bc=1) are supported")
    if tpp != 0.0:
        raise NotImplementedError("tpp != 0 is not supported by the bond maps")
    cplx = nflux != 0 or twistx != 0 or twisty != 0
    dnum = np.complex128 if cplx else np.float64
    if init_rng is None:
        raise ValueError("init_rng is required (use splitmix64_seed(seed))")
    init_rng = [int(v) for v in init_rng]
    rng = XorShift1024Star(init_rng)
    if geometry not in NORB:
        raise NotImplementedError("geometry %r" % geometry)
    N = NORB[geometry] * Nx * Ny
    hs = init_hs_field(rng, L, N)

    if tog["meas_gen_suscept"] and not trans_sym:
        raise ValueError("Can't measure generalized susceptibility without trans_sym")
    if geometry == "square":
        maps = _square_maps(Nx, Ny, tp, trans_sym)
        kij, peierls = hopping_square(Nx, Ny, t=1.0, tp=tp, tpp=tpp, nflux=nflux, alpha=alpha,
                                      twistx=twistx, twisty=twisty)
        thermal = _thermal_phases(maps, peierls, tp, N)
    else:
        # the reference builds no bonds for these lattices (placeholders,
        # util/gen_1band_unified_hub.py:137-160, 697-903) and its tight-binding
        # models have no t' (util/tight_binding.py:195-196)
        if tp != 0.0:
            raise NotImplementedError("t' on the %s lattice" % geometry)
        if twistx != 0.0 or twisty != 0.0:
            raise NotImplementedError("twisted boundaries on the %s lattice" % geometry)
        for k in ("meas_bond_corr", "meas_thermal", "meas_2bond_corr", "meas_energy_corr",
                  "meas_nematic_corr", "meas_local_JQ", "meas_pair_bb_only"):
            if tog[k]:
                raise NotImplementedError("%s needs bonds, which the %s lattice does not define"
                                          % (k, geometry))
        maps = _orbital_maps(geometry, Nx, Ny, trans_sym)
        hop = {"triangular": hopping_triangular, "honeycomb": hopping_honeycomb,
               "kagome": hopping_kagome}[geometry]
        kij, peierls = hop(Nx, Ny, t=1.0, nflux=nflux, alpha=alpha)
        thermal = np.ones((maps["b2ps"], N), dtype=np.complex128)

    if not cplx:
        thermal = thermal.real
        kij = kij.real
        peierls = peierls.real
    Ku = kij.copy()
    Kd = Ku.copy()
    for i in range(N):
        Ku[i, i] -= mu - h
        Kd[i, i] -= mu + h

    num_i = maps["num_i"]
    U_i = U * np.ones(num_i, dtype=np.float64)
    exp_lmbd = np.exp(0.5 * U_i * dt) + np.sqrt(np.expm1(U_i * dt))
    mi = maps["map_i"]
    exp_lambda = np.array((exp_lmbd[mi] ** -1, exp_lmbd[mi]))
    dell = np.array((exp_lmbd[mi] ** 2 - 1, exp_lmbd[mi] ** -2 - 1))
    i32 = lambda v: np.array(v, dtype=np.int32)

    p = {
        "N": i32(N), "L": i32(L), "Nx": i32(Nx), "Ny": i32(Ny),
        "map_i": maps["map_i"], "map_ij": maps["map_ij"], "bonds": maps["bonds"],
        "bond2s": maps["bond2s"], "plaqs": maps["plaqs"], "map_plaq": maps["map_plaq"],
        "map_b": maps["map_b"], "map_bs": maps["map_bs"], "map_bb": maps["map_bb"],
        "map_b2": maps["map_b2"], "map_b2b": maps["map_b2b"], "map_bb2": maps["map_bb2"],
        "map_b2b2": maps["map_b2b2"],
        "peierlsu": peierls, "peierlsd": peierls.copy(),
        "pp_u": thermal.conj(), "pp_d": thermal.conj(), "ppr_u": thermal, "ppr_d": thermal.copy(),
        "Ku": Ku, "Kd": Kd, "U": U_i, "dt": np.array(dt, dtype=np.float64),
        "n_matmul": i32(n_matmul), "n_delay": i32(n_delay),
        "n_sweep_warm": i32(n_sweep_warm), "n_sweep_meas": i32(n_sweep_meas),
        "period_eqlt": i32(period_eqlt), "period_uneqlt": i32(period_uneqlt),
    }
    for k in ("meas_bond_corr", "meas_thermal", "meas_2bond_corr", "meas_energy_corr",
              "meas_nematic_corr", "meas_chiral", "meas_local_JQ", "meas_pair_bb_only",
              "meas_gen_suscept"):
        p[k] = np.array(tog[k], dtype=np.int64)
    p["checkpoint_every"] = np.array(checkpoint_every, dtype=np.int64)
    for k in ("num_i", "num_ij", "num_plaq_accum", "num_b_accum", "num_b2_accum"):
        p[k] = np.array(maps[k], dtype=np.int64)
    p["num_plaq"] = np.array(maps["num_plaq"], dtype=np.int64)
    for k in ("num_b", "num_b2", "num_bs", "num_bb", "num_b2b", "num_bb2", "num_b2b2"):
        p[k] = np.array(maps[k], dtype=np.int64)
    for k in ("i", "ij", "plaq", "b", "b2", "bs", "bb", "bb2", "b2b", "b2b2"):
        p["degen_" + k] = np.array(maps["degen_" + k][0], dtype=np.int32)
    p["exp_Ku"] = expm(-dt * Ku)
    p["exp_Kd"] = expm(-dt * Kd)
    p["inv_exp_Ku"] = expm(dt * Ku)
    p["inv_exp_Kd"] = expm(dt * Kd)
    p["exp_halfKu"] = expm(-dt / 2 * Ku)
    p["exp_halfKd"] = expm(-dt / 2 * Kd)
    p["inv_exp_halfKu"] = expm(dt / 2 * Ku)
    p["inv_exp_halfKd"] = expm(dt / 2 * Kd)
    p["exp_lambda"] = exp_lambda
    p["del"] = dell
    p["F"] = i32(L // n_matmul)
    p["n_sweep"] = i32(n_sweep_warm + n_sweep_meas)

    state = {"sweep": i32(0), "init_rng": np.array(init_rng, dtype=np.uint64),
             "rng": rng.state(), "hs": hs, "partial_write": np.array(0, dtype=np.int64)}
    z = lambda n: np.zeros(n, dtype=dnum)
    num_ij = maps["num_ij"]
    meq = {"n_sample": i32(0), "sign": np.array(0.0, dtype=dnum)}
    for k in ("density", "density_u", "density_d", "double_occ"):
        meq[k] = z(num_i)
    for k in ("g00", "g00_u", "g00_d", "nn", "xx", "zz", "pair_sw"):
        meq[k] = z(num_ij)
    if tog["meas_gen_suscept"]:
        for k in ("uuuu", "dddd", "dduu", "uudd"):
            meq[k] = z(num_ij * num_ij)
    if tog["meas_chiral"]:
        meq["chi"] = z(maps["num_plaq_accum"])
    if tog["meas_energy_corr"]:
        meq["kk"] = z(maps["num_bb"])
        meq["kv"] = z(maps["num_bs"])
        meq["kn"] = z(maps["num_bs"])
        meq["vv"] = z(num_ij)
        meq["vn"] = z(num_ij)
    if tog["meas_local_JQ"]:
        assert trans_sym == 0
        meq["j"] = z(maps["num_b"])
        meq["jn"] = z(maps["num_b"])
        meq["j2"] = z(maps["num_b2"])
    mue = {}
    if period_uneqlt > 0:
        mue = {"n_sample": i32(0), "sign": np.array(0.0, dtype=dnum)}
        for k in ("gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw"):
            mue[k] = z(num_ij * L)
        if tog["meas_gen_suscept"]:
            for k in ("uuuu", "dddd", "dduu", "uudd"):
                mue[k] = z(num_ij * num_ij * L)
        nbb, nb2b, nbb2, nb2b2 = (maps["num_bb"], maps["num_b2b"], maps["num_bb2"],
                                  maps["num_b2b2"])
        if tog["meas_pair_bb_only"]:
            assert not any(tog[k] for k in ("meas_thermal", "meas_bond_corr", "meas_2bond_corr",
                                            "meas_energy_corr", "meas_nematic_corr",
                                            "meas_gen_suscept"))
            assert not trans_sym
            mue["pair_bb"] = z(nbb * L)
        if tog["meas_bond_corr"]:
            for k in ("pair_bb", "jj", "jsjs", "kk", "ksks"):
                mue[k] = z(nbb * L)
        if tog["meas_thermal"]:
            mue["j2jn"] = z(nb2b * L)
            mue["jnj2"] = z(nbb2 * L)
            for k in ("jnjn", "jjn", "jnj"):
                mue[k] = z(nbb * L)
        if tog["meas_2bond_corr"]:
            mue["j2j2"] = z(nb2b2 * L)
            mue["j2j"] = z(nb2b * L)
            mue["jj2"] = z(nbb2 * L)
            for k in ("pair_b2b2", "js2js2", "k2k2", "ks2ks2"):
                mue[k] = z(nb2b2 * L)
        if tog["meas_energy_corr"]:
            mue["kv"] = z(maps["num_bs"] * L)
            mue["kn"] = z(maps["num_bs"] * L)
            mue["vv"] = z(num_ij * L)
            mue["vn"] = z(num_ij * L)
        if tog["meas_nematic_corr"]:
            mue["nem_nnnn"] = z(nbb * L)
            mue["nem_ssss"] = z(nbb * L)
    meta = {"model": "Hubbard (complex)" if cplx else "Hubbard", "Nx": Nx, "Ny": Ny,
            "Norb": NORB[geometry], "bps": maps["bps"], "b2ps": maps["b2ps"],
            "plaq_per_cell": maps["plaq_per_cell"], "U": U, "t'": tp,
            "t''": tpp, "nflux": nflux, "twistx": twistx, "twisty": twisty, "mu": mu, "h": h,
            "beta": L * dt, "trans_sym": trans_sym, "geometry": geometry, "bc": bc,
            "gauge_alpha": alpha,
            "commit": "dqmc-mi355x", "version": 0.1}
    return {"params": p, "state": state, "meas_eqlt": meq, "meas_uneqlt": mue, "metadata": meta}


def create_batch(Nfiles=1, seed=None, **kwargs):
    """List of Nfiles simulations sharing params; file i>0 starts from the seed
    state advanced by i jumps (gen_1band_unified_hub.py:1163-1195)."""
    if seed is None:
        raise ValueError("seed is required")
    rng0 = XorShift1024Star(splitmix64_seed(seed))
    sims = [create_1(init_rng=list(rng0.s), **kwargs)]
    L = int(sims[0]["params"]["L"])
    N = int(sims[0]["params"]["N"])
    for _ in range(1, Nfiles):
        rng0.jump()
        init = list(rng0.s)
        r = XorShift1024Star(init)
        hs = init_hs_field(r, L, N)
        s = {"params": sims[0]["params"], "metadata": sims[0]["metadata"],
             "state": dict(sims[0]["state"]),
             "meas_eqlt": {k: v.copy() for k, v in sims[0]["meas_eqlt"].items()},
             "meas_uneqlt": {k: v.copy() for k, v in sims[0]["meas_uneqlt"].items()}}
        s["state"]["init_rng"] = np.array(init, dtype=np.uint64)
        s["state"]["rng"] = r.state()
        s["state"]["hs"] = hs
        sims.append(s)
    return sims

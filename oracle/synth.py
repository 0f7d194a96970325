"""Counter-based synthetic spin-orbital integrals (test infrastructure only).

The synthetic workloads are not present in the reference (it only consumes RHF
integrals through ``Hartree_Fock_Solver``, ccsd_memory_optimized.cpp:49-74).
The generator is defined here and restated bit-for-bit in the HIP library's
on-device generator:

* ``h(seed, c) = mix64(seed + (c + 1) * 0x9E3779B97F4A7C15)`` -- the c-th output
  of a SplitMix64 stream started at ``seed`` (Steele, Lea & Flood 2014).
* orbital energies: ``u_p = (h(seed ^ 0xD1B54A32D192ED03, p) >> 11) * 2**-53``;
  occupied ``-1.5 + 1.2 u``, virtual ``0.2 + 2.8 u``, each group sorted ascending;
  Fock matrix ``f = diag(eps)``.
* integrals: every canonical key ``K`` (one per class of <pq||rs> under
  antisymmetry and pq<->rs symmetry) gets ``0.02 * N(0,1)`` by Box-Muller
  (Box & Muller 1958) on ``u1 = ((h(seed,2K) >> 11) + 1) * 2**-53`` and
  ``u2 = (h(seed,2K+1) >> 11) * 2**-53``: ``g = 0.02*sqrt(-2 ln u1)*cos(2 pi u2)``.
* canonical key: orbitals are numbered occupied first (0..O-1) then virtual
  (O..O+V-1); ``(p,q)`` and ``(r,s)`` are sorted (each swap flips the sign, a
  repeated index gives 0), pair index ``P = q(q-1)/2 + p``; ``K = hi(hi+1)/2 + lo``
  with ``lo, hi = sorted(P, R)``.

Packed block layouts (row-major, pair index ``pidx(p,q) = q(q-1)/2 + p`` for
p<q) are the library's input format:

=======  =====================  ==========================
block    shape                  entry
=======  =====================  ==========================
oooo     P2(O) x P2(O)          <ij||kl>, i<j, k<l
ooov     P2(O) x (O*V)          <ij||ka>, i<j, col k*V+a
oovv     P2(O) x P2(V)          <ij||ab>, i<j, a<b
ovov     (O*V) x (O*V)          <ia||jb>, row i*V+a
ovvv     (O*V) x P2(V)          <ia||bc>, b<c
vvvv     P2(V) x P2(V)          <ab||cd>, a<b, c<d
=======  =====================  ==========================
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_EPS_SALT = 0xD1B54A32D192ED03
_TWO53 = 2.0 ** -53

BLOCKS = ("oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv")


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def hstream(seed, c):
    """c-th output of a SplitMix64 stream seeded with ``seed`` (vectorised)."""
    c = np.asarray(c, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) + (c + np.uint64(1)) * _GAMMA)


def p2(n):
    return n * (n - 1) // 2


def pair_list(n):
    """(p, q) arrays with p<q, ordered by pair index q(q-1)/2 + p."""
    q = np.concatenate([np.full(k, k, dtype=np.int64) for k in range(n)]) if n > 1 else np.zeros(0, np.int64)
    p = np.concatenate([np.arange(k, dtype=np.int64) for k in range(n)]) if n > 1 else np.zeros(0, np.int64)
    return p, q


def pidx(p, q):
    return q * (q - 1) // 2 + p


def orbital_energies(O, V, seed):
    s = np.uint64(np.uint64(seed) ^ np.uint64(_EPS_SALT))
    u = (hstream(s, np.arange(O + V)) >> np.uint64(11)).astype(np.float64) * _TWO53
    eo = np.sort(-1.5 + 1.2 * u[:O])
    ev = np.sort(0.2 + 2.8 * u[O:])
    return eo, ev


def integral_values(p, q, r, s, seed, scale=0.02):
    """<pq||rs> for global spin-orbital indices (arrays, broadcast)."""
    p, q, r, s = (np.asarray(x, dtype=np.int64) for x in np.broadcast_arrays(p, q, r, s))
    sign = np.ones(p.shape, dtype=np.float64)
    sw = p > q
    p, q = np.where(sw, q, p), np.where(sw, p, q)
    sign = np.where(sw, -sign, sign)
    sw = r > s
    r, s = np.where(sw, s, r), np.where(sw, r, s)
    sign = np.where(sw, -sign, sign)
    sign = np.where((p == q) | (r == s), 0.0, sign)
    P = pidx(p, q)
    R = pidx(r, s)
    lo = np.minimum(P, R).astype(np.uint64)
    hi = np.maximum(P, R).astype(np.uint64)
    K = hi * (hi + np.uint64(1)) // np.uint64(2) + lo
    two = np.uint64(2)
    u1 = ((hstream(seed, two * K) >> np.uint64(11)).astype(np.float64) + 1.0) * _TWO53
    u2 = (hstream(seed, two * K + np.uint64(1)) >> np.uint64(11)).astype(np.float64) * _TWO53
    g = scale * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return sign * g


def _index_sets(O, V):
    io, jo = pair_list(O)
    av, bv = pair_list(V)
    ov_i = np.repeat(np.arange(O), V)
    ov_a = np.tile(np.arange(V), O) + O
    return dict(oo=(io, jo), vv=(av + O, bv + O), ov=(ov_i, ov_a))


_BLOCK_KINDS = {
    "oooo": ("oo", "oo"),
    "ooov": ("oo", "ov"),
    "oovv": ("oo", "vv"),
    "ovov": ("ov", "ov"),
    "ovvv": ("ov", "vv"),
    "vvvv": ("vv", "vv"),
}


def packed_block(name, O, V, seed, scale=0.02, rows=None, dtype=np.float64, workers=1):
    """Generate one packed block (optionally only a row range ``rows=(r0, r1)``);
    ``workers`` > 1 fills row chunks on that many threads (numpy releases the
    GIL inside the element-wise kernels)."""
    sets = _index_sets(O, V)
    rk, ck = _BLOCK_KINDS[name]
    rp, rq = sets[rk]
    cp, cq = sets[ck]
    if rows is not None:
        rp, rq = rp[rows[0]:rows[1]], rq[rows[0]:rows[1]]
    out = np.empty((len(rp), len(cp)), dtype=dtype)
    chunk = max(1, 4_000_000 // max(1, len(cp)))

    def fill(r0):
        r1 = min(len(rp), r0 + chunk)
        out[r0:r1] = integral_values(rp[r0:r1, None], rq[r0:r1, None], cp[None, :], cq[None, :], seed, scale)

    starts = range(0, len(rp), chunk)
    if workers > 1 and len(starts) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(fill, starts))
    else:
        for r0 in starts:
            fill(r0)
    return out


def make_problem(O, V, seed, scale=0.02, blocks=BLOCKS, workers=1):
    """Synthetic problem: Fock matrix (diag) and packed integral blocks."""
    eo, ev = orbital_energies(O, V, seed)
    f = np.diag(np.concatenate([eo, ev]))
    prob = dict(O=O, V=V, seed=seed, f=f)
    for b in blocks:
        prob[b] = packed_block(b, O, V, seed, scale, workers=workers)
    return prob


# ---------------------------------------------------------------------------
# packed <-> dense helpers used by the dense oracles
# ---------------------------------------------------------------------------

def unpack_pairs(x, n, axis):
    """Expand a packed (p<q) axis of length P2(n) into two antisymmetric axes (n, n)."""
    x = np.moveaxis(np.asarray(x), axis, -1)
    p, q = pair_list(n)
    out = np.zeros(x.shape[:-1] + (n, n), dtype=x.dtype)
    out[..., p, q] = x
    out[..., q, p] = -x
    return np.moveaxis(out, [-2, -1], [axis, axis + 1])


def pack_pairs(x, n, axis):
    """Inverse of unpack_pairs (takes the p<q entries)."""
    x = np.moveaxis(np.asarray(x), [axis, axis + 1], [-2, -1])
    p, q = pair_list(n)
    return np.moveaxis(x[..., p, q], -1, axis)


def dense_blocks(prob):
    """Dense antisymmetric blocks from a packed problem."""
    O, V = prob["O"], prob["V"]
    d = {}
    d["oooo"] = unpack_pairs(unpack_pairs(prob["oooo"], O, 1), O, 0)
    d["ooov"] = unpack_pairs(prob["ooov"].reshape(p2(O), O, V), O, 0)
    d["oovv"] = unpack_pairs(unpack_pairs(prob["oovv"], V, 1), O, 0)
    d["ovov"] = prob["ovov"].reshape(O, V, O, V)
    d["ovvv"] = unpack_pairs(prob["ovvv"], V, 1).reshape(O, V, V, V)
    d["vvvv"] = unpack_pairs(unpack_pairs(prob["vvvv"], V, 1), V, 0)
    return d


def full_tensor(O, V, seed, scale=0.02):
    """Full dense <pq||rs> over all N=O+V spin orbitals (tiny sizes only)."""
    n = O + V
    idx = np.arange(n)
    return integral_values(idx[:, None, None, None], idx[None, :, None, None],
                           idx[None, None, :, None], idx[None, None, None, :], seed, scale)


def blocks_from_full(g, O):
    """Dense blocks sliced from a full <pq||rs> tensor."""
    o, v = slice(0, O), slice(O, None)
    return dict(oooo=g[o, o, o, o], ooov=g[o, o, o, v], oovv=g[o, o, v, v],
                ovov=g[o, v, o, v], ovvv=g[o, v, v, v], vvvv=g[v, v, v, v])


def pack_dense(d, O, V):
    """Packed blocks from dense antisymmetric blocks (inverse of dense_blocks)."""
    out = {}
    out["oooo"] = pack_pairs(pack_pairs(d["oooo"], O, 0), O, 1)
    out["ooov"] = pack_pairs(d["ooov"], O, 0).reshape(p2(O), O * V)
    out["oovv"] = pack_pairs(pack_pairs(d["oovv"], O, 0), V, 1)
    out["ovov"] = np.ascontiguousarray(d["ovov"]).reshape(O * V, O * V)
    out["ovvv"] = pack_pairs(np.asarray(d["ovvv"]).reshape(O * V, V, V), V, 1)
    out["vvvv"] = pack_pairs(pack_pairs(d["vvvv"], V, 0), V, 1)
    return {k: np.ascontiguousarray(x) for k, x in out.items()}

// lekf_regional.hip -- regional (SCALE-style) obs index and capped selection (gfx950).
//
// Obs index (scale/letkf/letkf_obs.f90:740-830, one subdomain): obs go to
// cell (ceiling(ri-0.5), ceiling(rj-0.5)) clamped to the grid, keeping input
// order within a cell (stable sort of unique keys cell << 32 | index).
//
// Selection (scale/letkf/letkf_tools.f90:853-1178): the rectangle of cells
// floor(ri -/+ dlon_zero) .. ceiling(...) is scanned row by row; a candidate
// passes when ndist = (sqrt(rdx^2+rdy^2)/sigma_h)^2 + (|dz|/sigma_v)^2 is not
// above the literal 13.33333333, with rloc = exp(-ndist/2) and
// rdiag = err^2/rloc.  Without a cap the accepted obs are emitted in scan
// order.
//
// The cap MAX_NOBS_PER_GRID = nobs_max > 0 (:1041-1156).  Per bucket (the
// reference's (element, obs type) pair) the reference keeps an
// insertion-sorted set of at most nobs_max obs by rdiag in storage slots: an
// obs that arrives while the set is not full takes the next free slot; later,
// an obs with rdiag strictly below the set's last-priority entry evicts that
// entry and takes over its slot (so an earlier obs wins a tie).  The kept obs
// are emitted bucket by bucket in slot order.  cap_slots gives the same slots
// without inserting serially:
//   * a candidate enters when fewer than nobs_max candidates before it in scan
//     order precede it in (rdiag, scan position) order, and the set after any
//     prefix is the nobs_max first entries of that prefix in this order;
//   * the set's last entry only moves down, so entries are evicted from the
//     largest down: the i-th candidate that enters a full set evicts the i-th
//     largest of the entries ever evicted, and takes its slot.
// A block of 256 candidates at a time is merged into the set held in LDS
// (ascending, double-buffered): each candidate's count of preceding entries
// is a binary search in the set plus a count over the block's earlier
// candidates, the new set and the block's evictions come from ranks in the
// union, and a candidate evicted within its own block passes its slot on
// (pointer jumping).
#include "lekf_dev.hpp"
#include "lekf_regional.hpp"
#include "lekf_tune.hpp"
#include "lekf_ktime.hpp"

// Every FP64 operation in this file is rounded as written: no FMA contraction
// (the reference's results are reproduced operation by operation).
#pragma clang fp contract(off)

namespace {

constexpr double kNdMax = 13.33333333;  // scale/letkf/letkf_obs.f90:30

__device__ __forceinline__ unsigned long long lane_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

constexpr int kSelT = 256;  // threads per workgroup of the selection (4 waves)
constexpr int kSelW = kSelT / 64;

constexpr int kCapU = 1;               // candidates per thread in a block of the cap
constexpr int kCapB = kSelT * kCapU;   // candidates per block

// One block's potential entrants, entrants and evictions of the workgroup
// form of the cap (caps above LEKF_CAP_LDS, the kept set S in global memory).
struct CapLds {
    unsigned long long pkey[kCapB];  // potential entrants, scan order
    int ppos[kCapB];
    unsigned long long nkey[kCapB];  // entrants, scan order
    int npos[kCapB];
    int nc1[kCapB];                  // entries of S before each entrant
    int nU[kCapB];                   // entrant's rank in S + entrants
    unsigned long long ekey[kCapB];  // entrants, ascending
    int epos[kCapB];
    int ev[kCapB];                   // evictions, ascending: slot (>= 0) or -(entrant + 1)
    int slotn[kCapB];
    int link[kCapB];
    int wc[2][kCapU][kSelW];
};

// S in global memory (caps above LEKF_CAP_LDS): 2 * cap entries per array
struct CapGlobal {
    unsigned long long* key;
    int* pos;
    int* slot;
};

// One wave's cap (caps up to LEKF_CAP_LDS): the kept set S ascending by
// (key, candidate) with each entry's storage slot, a queue of potential
// entrants in scan order, and one block's entrant c1 values, evictions and
// slot links.
struct CapEnt {
    unsigned long long key;
    int pos;
    int slot;
};

struct WaveCapLds {
    CapEnt s[LEKF_CAP_LDS];
    unsigned long long qkey[128];
    int qpos[128];
    int c1l[64];
    int ev[64];
    int slotn[64];
    int link[64];
};

constexpr int kWaveCaps = 2;  // waves of a selection workgroup that cap bucket streams

struct SelLds {
    union {
        CapLds cap;
        WaveCapLds wcap[kWaveCaps];
    };
    int seg_beg[LEKF_REG_SEG];
    int seg_pre[LEKF_REG_SEG + 1];
    int wcnt[kSelW][LEKF_REG_NB];
};

struct Cand {
    bool ok;
    int o, b;
    double rho, rd;
    unsigned long long bits;
};

struct ObsRec {
    double ri, rj, z, err;
    int b;
};

__device__ __forceinline__ ObsRec load_obs(const RegionalArgs& A, int o)
{
    ObsRec r;
    r.ri = A.ori[o];
    r.rj = A.orj[o];
    r.z = A.oz[o];
    r.err = A.oerr[o];
    r.b = A.obkt[o];
    return r;
}

// The reference's test and weights (scale/letkf/letkf_tools.f90:978-1037),
// each operation rounded as written (no contraction), so the cutoff decision
// and rdiag are bit-identical to the reference's.
__device__ __forceinline__ Cand eval_rec(const RegionalArgs& A, const ObsRec& r, int o, double ri,
                                         double rj, double rz)
{
    Cand c;
    const double rdx = rn_mul(rn_sub(ri, r.ri), A.dx);
    const double rdy = rn_mul(rn_sub(rj, r.rj), A.dy);
    const double nd_h = rn_div(__dsqrt_rn(rn_add(rn_mul(rdx, rdx), rn_mul(rdy, rdy))),
                                  A.sigma_h);
    const double nd_v = rn_div(fabs(rn_sub(r.z, rz)), A.sigma_v);
    const double nd = rn_add(rn_mul(nd_h, nd_h), rn_mul(nd_v, nd_v));
    c.o = o;
    c.ok = !(nd > kNdMax);
    c.b = r.b;
    c.rho = 0.0;
    c.rd = 0.0;
    c.bits = 0;
    if (c.ok) {
        c.rho = exp(-0.5 * nd);
        c.rd = rn_div(rn_mul(r.err, r.err), c.rho);
        c.bits = (unsigned long long)__double_as_longlong(c.rd);
    }
    return c;
}

// candidate q of the point's scan (row segments in LDS)
__device__ __forceinline__ int cand_obs(const SelLds& S, int nseg, int q, int& seg)
{
    while (seg + 1 < nseg && S.seg_pre[seg + 1] <= q) ++seg;
    return S.seg_beg[seg] + (q - S.seg_pre[seg]);
}

// Block-wide compaction of U flags per thread in (u, thread) order: pos[u]
// is the position of entry u among the flagged ones, tot their number.  The
// counts alternate between two buffers, so one barrier suffices.  Uniform
// call.
template <int U>
__device__ __forceinline__ void block_rank(int (*wc)[U][kSelW], int& par, const bool (&flag)[U],
                                           int tid, int (&pos)[U], int& tot)
{
    const int lane = tid & 63, wv = tid >> 6;
    unsigned long long m[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        m[u] = __ballot(flag[u]);
        if (lane == 0) wc[par][u][wv] = __popcll(m[u]);
    }
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int off = 0, t = 0;
#pragma unroll
        for (int w = 0; w < kSelW; ++w) {
            const int c = wc[par][u][w];
            off += w < wv ? c : 0;
            t += c;
        }
        pos[u] = base + off + __popcll(m[u] & lane_lt(lane));
        base += t;
    }
    tot = base;
    par ^= 1;
}

// V binary searches in lockstep (their loads overlap): out[v] = entries of the
// ascending (key, candidate) array [0, n) before (x[v], xp[v])
template <int V>
__device__ __forceinline__ void count_before(const unsigned long long* key, const int* pos, int n,
                                             const unsigned long long (&x)[V],
                                             const int (&xp)[V], int (&out)[V])
{
    int lo[V], hi[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        lo[v] = 0;
        hi[v] = n;
    }
    for (int w = n; w > 0; w >>= 1) {
        unsigned long long y[V];
        int yp[V], mid[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            mid[v] = (lo[v] + hi[v]) >> 1;
            const int m = mid[v] < n ? mid[v] : 0;
            y[v] = key[m];
            yp[v] = pos[m];
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (lo[v] < hi[v]) {
                if (y[v] < x[v] || (y[v] == x[v] && yp[v] < xp[v]))
                    lo[v] = mid[v] + 1;
                else
                    hi[v] = mid[v];
            }
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = lo[v];
}

// The reference's cap (scale/letkf/letkf_tools.f90:1041-1156) over one
// bucket's candidates 0..n-1 in scan order, keyat(i) the priority key of
// candidate i (smaller first), by all kSelT threads of a workgroup with the
// kept set in global memory (caps above LEKF_CAP_LDS).  Returns the kept
// count; the kept entries are G[buf][0..kept) with pos = candidate and
// slot = storage slot (the output position).  Uniform call.
template <class KeyAt>
__device__ int cap_slots(CapLds& L, const CapGlobal& G, KeyAt keyat, int n, int cap, int tid,
                         int& bufo)
{
    constexpr bool kLds = false;
    constexpr int U = kCapU;
    constexpr int V = LEKF_CAP_LDS / kSelT;  // S entries per thread per round of the merge
    int s = 0, buf = 0, par = 0;
    unsigned long long xn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = u * kSelT + tid;
        xn[u] = i < n ? keyat(i) : 0ull;
    }
    for (int b0 = 0; b0 < n; b0 += kCapB) {
        (void)kLds;
        unsigned long long* Sk = G.key + (size_t)buf * cap;
        int* Sp = G.pos + (size_t)buf * cap;
        int* Ss = G.slot + (size_t)buf * cap;
        unsigned long long* Tk = G.key + (size_t)(buf ^ 1) * cap;
        int* Tp = G.pos + (size_t)(buf ^ 1) * cap;
        int* Ts = G.slot + (size_t)(buf ^ 1) * cap;
        unsigned long long x[U];
        bool pot[U];
        const unsigned long long tau = s == cap ? Sk[cap - 1] : 0ull;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x[u] = xn[u];
            const int i = b0 + u * kSelT + tid;
            // into a full set only below its last entry (a later candidate
            // with an equal key comes after it)
            pot[u] = i < n && (s < cap || x[u] < tau);
            const int inext = i + kCapB;  // the next block's keys, loaded ahead
            xn[u] = inext < n ? keyat(inext) : 0ull;
        }
        int pr[U], np;
        block_rank<U>(L.wc, par, pot, tid, pr, np);
        if (np == 0) continue;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (pot[u]) {
                L.pkey[pr[u]] = x[u];
                L.ppos[pr[u]] = b0 + u * kSelT + tid;
            }
        __syncthreads();
        // potential q = u * kSelT + tid of this thread: its entries of S
        // before it (c1) and the earlier potentials not above it (c2)
        unsigned long long y[U];
        int yp[U], c1[U], c2[U];
        int jmax = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = u * kSelT + tid;
            y[u] = q < np ? L.pkey[q] : ~0ull;
            yp[u] = q < np ? L.ppos[q] : 0;
            c2[u] = 0;
            if (q < np) jmax = q;
        }
        count_before<U>(Sk, Sp, s, y, yp, c1);
        {
            int j = 0;
            for (; j + 4 <= jmax; j += 4) {
                const unsigned long long k0 = L.pkey[j], k1 = L.pkey[j + 1];
                const unsigned long long k2 = L.pkey[j + 2], k3 = L.pkey[j + 3];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = u * kSelT + tid;
                    c2[u] += (j < q && k0 <= y[u]) + (j + 1 < q && k1 <= y[u]) +
                             (j + 2 < q && k2 <= y[u]) + (j + 3 < q && k3 <= y[u]);
                }
            }
            for (; j < jmax; ++j) {
                const unsigned long long k0 = L.pkey[j];
#pragma unroll
                for (int u = 0; u < U; ++u) c2[u] += (j < u * kSelT + tid && k0 <= y[u]);
            }
        }
        bool ent[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ent[u] = u * kSelT + tid < np && c1[u] + c2[u] < cap;
        int er[U], e;
        block_rank<U>(L.wc, par, ent, tid, er, e);
        if (e == 0) continue;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ent[u]) {
                L.nkey[er[u]] = y[u];
                L.npos[er[u]] = yp[u];
                L.nc1[er[u]] = c1[u];
            }
        __syncthreads();
        // entrant t = u * kSelT + tid: its rank among the entrants
        {
            unsigned long long xk[U];
            int r[U];
            int tmax = -1;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = u * kSelT + tid;
                xk[u] = t < e ? L.nkey[t] : 0ull;
                r[u] = 0;
                if (t < e) tmax = t;
            }
            if (tmax >= 0) {
                int w = 0;
                for (; w + 4 <= e; w += 4) {
                    const unsigned long long z0 = L.nkey[w], z1 = L.nkey[w + 1];
                    const unsigned long long z2 = L.nkey[w + 2], z3 = L.nkey[w + 3];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int t = u * kSelT + tid;
                        r[u] += (z0 < xk[u] || (z0 == xk[u] && w < t)) +
                                (z1 < xk[u] || (z1 == xk[u] && w + 1 < t)) +
                                (z2 < xk[u] || (z2 == xk[u] && w + 2 < t)) +
                                (z3 < xk[u] || (z3 == xk[u] && w + 3 < t));
                    }
                }
                for (; w < e; ++w) {
                    const unsigned long long z0 = L.nkey[w];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        r[u] += (z0 < xk[u] || (z0 == xk[u] && w < u * kSelT + tid));
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = u * kSelT + tid;
                if (t < e) {
                    L.ekey[r[u]] = xk[u];
                    L.epos[r[u]] = L.npos[t];
                    L.nU[t] = L.nc1[t] + r[u];
                }
            }
        }
        __syncthreads();
        const int snew = s + e < cap ? s + e : cap;
        const int nev = s + e - snew;
        // S entries move up by the entrants before them; the top nev leave
        for (int r0 = 0; r0 < s; r0 += V * kSelT) {
            unsigned long long xk[V];
            int xp[V], U2[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int r = r0 + v * kSelT + tid;
                xk[v] = r < s ? Sk[r] : ~0ull;
                xp[v] = r < s ? Sp[r] : 0x7fffffff;
            }
            count_before<V>(L.ekey, L.epos, e, xk, xp, U2);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int r = r0 + v * kSelT + tid;
                if (r < s) {
                    const int Ur = r + U2[v];
                    if (Ur < cap) {
                        Tk[Ur] = xk[v];
                        Tp[Ur] = xp[v];
                        Ts[Ur] = Ss[r];
                    } else {
                        L.ev[Ur - cap] = Ss[r];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = u * kSelT + tid;
            if (t < e) {
                const int Ut = L.nU[t];
                if (Ut < cap) {
                    Tk[Ut] = L.nkey[t];
                    Tp[Ut] = L.npos[t];
                } else {
                    L.ev[Ut - cap] = -(t + 1);
                }
            }
        }
        __syncthreads();
        // slots: the next free one while the set fills, then the slot of the
        // evicted entry, evictions taken from the largest down
        const int nfill = cap - s;
        bool chain = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = u * kSelT + tid;
            if (t < e) {
                int sl = -1, ln = -1;
                if (t < nfill) {
                    sl = s + t;
                } else {
                    const int v = L.ev[nev - 1 - (t - nfill)];
                    if (v >= 0)
                        sl = v;
                    else
                        ln = -v - 1;
                }
                L.slotn[t] = sl;
                L.link[t] = ln;
                chain |= sl < 0;
            }
        }
        // an entrant evicted within its block passes its slot on to the
        // entrant that evicted it: resolve the chains by pointer jumping (a
        // link points to an earlier entrant, so a chain of at most kCapB
        // entrants resolves within log2(kCapB) + 1 rounds; the bound only
        // guards the loop)
        if (__syncthreads_or(chain)) {
            for (int round = 0; round < 16; ++round) {
                bool un = false;
                int sl[U], ln[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = u * kSelT + tid;
                    sl[u] = -1;
                    ln[u] = -1;
                    if (t < e && L.slotn[t] < 0) {
                        un = true;
                        const int l = L.link[t];
                        sl[u] = L.slotn[l];
                        ln[u] = L.link[l];
                    }
                }
                if (!__syncthreads_or(un)) break;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = u * kSelT + tid;
                    if (t < e && L.slotn[t] < 0) {
                        if (sl[u] >= 0)
                            L.slotn[t] = sl[u];
                        else
                            L.link[t] = ln[u];
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = u * kSelT + tid;
            if (t < e && L.nU[t] < cap) Ts[L.nU[t]] = L.slotn[t];
        }
        __syncthreads();
        s = snew;
        buf ^= 1;
    }
    bufo = buf;
    return s;
}

// Cap one candidate stream (cap above LEKF_CAP_LDS) with the whole workgroup
// and write the kept entries in slot order: emit(slot, candidate).  Uniform
// call; ends with a barrier.
template <class KeyAt, class Emit>
__device__ int cap_stream(SelLds& S, const CapGlobal& G, KeyAt keyat, Emit emit, int n, int cap,
                          int tid)
{
    int buf = 0;
    const int kept = cap_slots(S.cap, G, keyat, n, cap, tid, buf);
    for (int r = tid; r < kept; r += kSelT) {
        const int sl = G.slot[(size_t)buf * cap + r];
        if ((unsigned)sl < (unsigned)kept) emit(sl, G.pos[(size_t)buf * cap + r]);
    }
    __syncthreads();
    return kept;
}

// ---------------------------------------------------------------------------
// The cap by one wave (caps up to LEKF_CAP_LDS).  The same merge as
// cap_slots, over blocks of 64 potential entrants: candidates after the
// first cap are filtered against the current last entry of S into a queue
// (the last entry only moves down, so a candidate filtered out can never
// enter, and a queued one is tested exactly later); each block of 64 queued
// candidates is sorted across the lanes by (key, candidate), which gives at
// once each one's count of earlier queued candidates not above it and the
// entrants' ranks; S is merged in place (every lane reads its 16 entries
// before any lane writes).  The first cap candidates fill S in scan order and
// are sorted by a bitonic network in LDS.
// ---------------------------------------------------------------------------

// order of the lanes' LDS traffic within the wave (no cross-wave sharing)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bool ent_less(unsigned long long ak, int ap, unsigned long long bk, int bp)
{
    return ak < bk || (ak == bk && ap < bp);
}

// entries of S[0, n) (ascending) before (x, xp)
__device__ __forceinline__ int wcount_before(const CapEnt* S, int n, unsigned long long x, int xp)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ent_less(S[mid].key, S[mid].pos, x, xp))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// ascending bitonic sort of the 64 lanes' (key, id) pairs (ids distinct)
__device__ __forceinline__ void wave_sort64(unsigned long long& k, int& id, int lane)
{
#pragma unroll
    for (int sz = 2; sz <= 64; sz <<= 1) {
#pragma unroll
        for (int j = sz >> 1; j > 0; j >>= 1) {
            const unsigned long long ok = __shfl_xor(k, j, 64);
            const int oid = __shfl_xor(id, j, 64);
            const bool up = (lane & sz) == 0;
            const bool lower = (lane & j) == 0;
            const bool oless = ent_less(ok, oid, k, id);
            if (lower == up ? oless : !oless) {
                k = ok;
                id = oid;
            }
        }
    }
}

// One block of np (<= 64) queued candidates against the full set S (cap
// entries): lane l takes queue entry l.  Returns nothing; S, its slots and
// its last key tau are updated.
__device__ void wave_cap_block(WaveCapLds& L, int np, int cap, int lane, unsigned long long& tau)
{
    const bool valid = lane < np;
    const unsigned long long y = valid ? L.qkey[lane] : ~0ull;
    const int yp = valid ? L.qpos[lane] : 0x7fffffff;
    const bool live = valid && y < tau;  // a queued candidate the threshold has passed stays out
    const int c1 = live ? wcount_before(L.s, cap, y, yp) : cap;
    // sort the live candidates by (key, lane): lane L holds the L-th
    unsigned long long sk = live ? y : ~0ull;
    int sid = lane;
    wave_sort64(sk, sid, lane);
    const bool slive = sk != ~0ull;
    // earlier (in scan order) live candidates sorted before this one
    unsigned long long pm = slive ? (1ull << sid) : 0ull;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long t = __shfl_up(pm, d, 64);
        if (lane >= d) pm |= t;
    }
    const unsigned long long before = pm & ~(slive ? (1ull << sid) : 0ull);
    const int c2 = __popcll(before & (sid == 0 ? 0ull : (~0ull >> (64 - sid))));
    const int c1s = __shfl(c1, sid, 64);
    const bool ents = slive && c1s + c2 < cap;
    const unsigned long long emS = __ballot(ents);
    const int e = __popcll(emS);
    if (e == 0) return;
    const int r = __popcll(emS & lane_lt(lane));  // rank among the entrants
    const int U = c1s + r;                        // position in S + entrants
    // scan-order index of each entrant: its flag back to the candidate's lane
    const int ento = __builtin_amdgcn_ds_permute(sid << 2, ents ? 1 : 0);
    const unsigned long long emO = __ballot(ento != 0);
    const int t = __shfl((int)__popcll(emO & lane_lt(lane)), sid, 64);
    if (ents) L.c1l[r] = c1s;
    wave_sync();
    // S entries move up by the entrants before them (entrant q precedes S[i]
    // when c1l[q] <= i); the top e leave S
    {
        constexpr int kPer = LEKF_CAP_LDS / 64;
        const int i0 = lane * kPer;
        unsigned long long vk[kPer];
        int vp[kPer], vs[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int i = i0 + u < cap ? i0 + u : 0;
            vk[u] = L.s[i].key;
            vp[u] = L.s[i].pos;
            vs[u] = L.s[i].slot;
        }
        int lo = 0, hi = e;  // entrants with c1 < i0
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (L.c1l[mid] < i0)
                lo = mid + 1;
            else
                hi = mid;
        }
        int ptr = lo;
        wave_sync();
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int i = i0 + u;
            if (i < cap) {
                while (ptr < e && L.c1l[ptr] <= i) ++ptr;
                const int ni = i + ptr;
                if (ni < cap) {
                    L.s[ni].key = vk[u];
                    L.s[ni].pos = vp[u];
                    L.s[ni].slot = vs[u];
                } else {
                    L.ev[ni - cap] = vs[u];
                }
            }
        }
    }
    if (ents) {
        if (U < cap) {
            CapEnt c;
            c.key = sk;
            c.pos = L.qpos[sid];
            c.slot = -1;
            L.s[U] = c;
        } else {
            L.ev[U - cap] = -(t + 1);
        }
    }
    wave_sync();
    // the i-th entrant in scan order evicts the i-th largest entry evicted
    // and takes its slot; an entrant evicted in its own block passes its slot
    // on to the entrant that evicted it (pointer jumping)
    int sl = -1, ln = -1;
    if (ents) {
        const int v = L.ev[e - 1 - t];
        if (v >= 0)
            sl = v;
        else
            ln = -v - 1;
        L.slotn[t] = sl;
        L.link[t] = ln;
    }
    wave_sync();
    for (int round = 0; round < 8 && __ballot(ents && sl < 0); ++round) {
        int s2 = -1, l2 = -1;
        if (ents && sl < 0) {
            s2 = L.slotn[ln];
            l2 = L.link[ln];
        }
        wave_sync();
        if (ents && sl < 0) {
            if (s2 >= 0)
                sl = s2;
            else
                ln = l2;
            L.slotn[t] = sl;
            L.link[t] = ln;
        }
        wave_sync();
    }
    if (ents && U < cap) L.s[U].slot = sl;
    wave_sync();
    tau = L.s[cap - 1].key;
}

// The reference's cap over one bucket's candidates 0..n-1 (n > cap) by one
// wave; then emit(slot, candidate) for the kept entries.  Returns cap.
template <class KeyAt, class Emit>
__device__ int wave_cap(WaveCapLds& L, KeyAt keyat, Emit emit, int n, int cap, int lane)
{
    // fill: the first cap candidates take slots 0..cap-1 in scan order
    int F = 64;
    while (F < cap) F <<= 1;
    for (int i = lane; i < F; i += 64) {
        CapEnt c;
        c.key = i < cap ? keyat(i) : ~0ull;
        c.pos = i < cap ? i : 0x7fffffff;
        c.slot = i;
        L.s[i] = c;
    }
    wave_sync();
    for (int sz = 2; sz <= F; sz <<= 1) {
        for (int j = sz >> 1; j > 0; j >>= 1) {
            for (int q = lane; q < F / 2; q += 64) {
                const int i = (q / j) * 2 * j + (q % j);
                const CapEnt a = L.s[i], bb = L.s[i + j];
                const bool up = (i & sz) == 0;
                if (ent_less(bb.key, bb.pos, a.key, a.pos) == up) {
                    L.s[i] = bb;
                    L.s[i + j] = a;
                }
            }
            wave_sync();
        }
    }
    unsigned long long tau = L.s[cap - 1].key;
    int q = 0;
    unsigned long long xn = cap + lane < n ? keyat(cap + lane) : 0ull;
    for (int b0 = cap; b0 < n; b0 += 64) {
        const unsigned long long x = xn;
        const int i = b0 + lane;
        xn = i + 64 < n ? keyat(i + 64) : 0ull;
        const bool pot = i < n && x < tau;
        const unsigned long long pm = __ballot(pot);
        if (pot) {
            const int o = q + __popcll(pm & lane_lt(lane));
            L.qkey[o] = x;
            L.qpos[o] = i;
        }
        q += __popcll(pm);
        wave_sync();
        if (q >= 64 || (b0 + 64 >= n && q > 0)) {
            const int nb = q < 64 ? q : 64;
            wave_cap_block(L, nb, cap, lane, tau);
            // the rest of the queue moves to its front
            unsigned long long k2 = 0;
            int p2 = 0;
            const bool mv = nb + lane < q;
            if (mv) {
                k2 = L.qkey[nb + lane];
                p2 = L.qpos[nb + lane];
            }
            wave_sync();
            if (mv) {
                L.qkey[lane] = k2;
                L.qpos[lane] = p2;
            }
            q -= nb;
            wave_sync();
            if (b0 + 64 >= n && q > 0) {
                wave_cap_block(L, q, cap, lane, tau);
                q = 0;
            }
        }
    }
    for (int r = lane; r < cap; r += 64) {
        const int sl = L.s[r].slot;
        if ((unsigned)sl < (unsigned)cap) emit(sl, L.s[r].pos);
    }
    wave_sync();
    return cap;
}

// Persistent selection: workgroup g handles points g, g + gridDim.x, ...
// Each candidate of a point is evaluated once.  Without a cap the accepted
// candidates are written in scan order.  With a cap each bucket's accepted
// candidates go, in scan order, to the bucket's stream in the workgroup's
// scratch (rdiag bits, obs, rho); a stream within the cap is written as it
// is, a longer one through cap_stream in slot order, bucket after bucket.
__global__ __launch_bounds__(kSelT) void reg_select(RegionalArgs A, long npt)
{
    __shared__ SelLds S;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int nij = A.nx * A.ny;
    const int nb = A.nbkt;
    const int cap = A.nobs_max;
    const size_t sbase = (size_t)blockIdx.x * (size_t)nb * (size_t)A.slot_cap;
    unsigned long long* skey = A.s_key + sbase;
    int* sob = A.s_ob + sbase;
    double* srho = A.s_rho + sbase;
    CapGlobal G;
    G.key = A.g_key ? A.g_key + (size_t)blockIdx.x * 2 * cap : nullptr;
    G.pos = A.g_pos ? A.g_pos + (size_t)blockIdx.x * 2 * cap : nullptr;
    G.slot = A.g_slot ? A.g_slot + (size_t)blockIdx.x * 2 * cap : nullptr;

    for (long pi = blockIdx.x; pi < npt; pi += gridDim.x) {
        const long b = A.pt_begin + pi;  // global point = lev * nij + ij
        const int lev = (int)(b / nij), ij = (int)(b % nij);
        const double ri = (double)(ij % A.nx + 1), rj = (double)(ij / A.nx + 1), rz = A.zlev[lev];
        int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
        int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
        if (imin < 1) imin = 1;
        if (imax > A.nx) imax = A.nx;
        if (jmin < 1) jmin = 1;
        if (jmax > A.ny) jmax = A.ny;
        const int nseg = (imin > imax || jmin > jmax) ? 0 : jmax - jmin + 1;
        if (tid == 0) {
            int acc = 0;
            for (int s = 0; s < nseg; ++s) {
                const int j = jmin + s;  // 1-based row
                // cellcum = inclusive cumulative count through cell (i, j), row-major
                const long rowbase = (long)(j - 1) * A.nx;
                const int beg = (imin > 1) ? A.cellcum[rowbase + imin - 2]
                                           : (rowbase > 0 ? A.cellcum[rowbase - 1] : 0);
                const int end = A.cellcum[rowbase + imax - 1];
                S.seg_beg[s] = beg;
                S.seg_pre[s] = acc;
                acc += end - beg;
            }
            S.seg_pre[nseg] = acc;
        }
        __syncthreads();
        const int ncand = S.seg_pre[nseg];
        const long obase = A.list_off[b] - A.list_base;  // relative to this chunk's list buffer

        // evaluate every candidate once; scan position q = q0 + tid
        int nacc = 0;
        int bcnt[LEKF_REG_NB];
#pragma unroll
        for (int t = 0; t < LEKF_REG_NB; ++t) bcnt[t] = 0;
        int seg = 0;
        for (int q0 = 0; q0 < ncand; q0 += kSelT) {
            const int q = q0 + tid;
            const int o = q < ncand ? cand_obs(S, nseg, q, seg) : -1;
            Cand c;
            c.ok = false;
            c.b = 0;
            if (o >= 0) c = eval_rec(A, load_obs(A, o), o, ri, rj, rz);
            unsigned long long m[LEKF_REG_NB];
#pragma unroll
            for (int t = 0; t < LEKF_REG_NB; ++t)
                m[t] = __ballot(c.ok && (cap <= 0 || c.b == t));
            if (lane == 0)
#pragma unroll
                for (int t = 0; t < LEKF_REG_NB; ++t) S.wcnt[wv][t] = __popcll(m[t]);
            __syncthreads();
            int pos = -1;
            if (cap <= 0) {
                int before = 0, tot = 0;
                for (int w = 0; w < kSelW; ++w) {
                    const int cw = S.wcnt[w][0];
                    before += w < wv ? cw : 0;
                    tot += cw;
                }
                pos = nacc + before + __popcll(m[0] & lane_lt(lane));
                nacc += tot;
            } else {
#pragma unroll
                for (int t = 0; t < LEKF_REG_NB; ++t) {
                    int before = 0, tot = 0;
                    for (int w = 0; w < kSelW; ++w) {
                        const int cw = S.wcnt[w][t];
                        before += w < wv ? cw : 0;
                        tot += cw;
                    }
                    if (c.ok && c.b == t) pos = bcnt[t] + before + __popcll(m[t] & lane_lt(lane));
                    bcnt[t] += tot;
                }
            }
            __syncthreads();
            if (c.ok) {
                if (cap <= 0) {
                    A.loc_idx[obase + pos] = c.o;
                    A.loc_rho[obase + pos] = c.rho;
                } else {
                    const size_t w = (size_t)c.b * (size_t)A.slot_cap + (size_t)pos;
                    skey[w] = c.bits;
                    sob[w] = c.o;
                    srho[w] = c.rho;
                }
            }
        }
        __syncthreads();
        if (cap > 0 && cap <= LEKF_CAP_LDS) {
            // wave w caps buckets w, w + kWaveCaps, ... (output offsets from the
            // counts, which every thread holds)
            nacc = 0;
            for (int t = 0; t < nb; ++t) {
                const int n = bcnt[t];
                if (wv < kWaveCaps && t % kWaveCaps == wv) {
                    const size_t so = (size_t)t * (size_t)A.slot_cap;
                    int* oi = A.loc_idx + obase + nacc;
                    double* orh = A.loc_rho + obase + nacc;
                    if (n <= cap) {
                        for (int i = lane; i < n; i += 64) {
                            oi[i] = sob[so + i];
                            orh[i] = srho[so + i];
                        }
                    } else {
                        wave_cap(
                            S.wcap[wv], [&](int i) { return skey[so + i]; },
                            [&](int slot, int i) {
                                oi[slot] = sob[so + i];
                                orh[slot] = srho[so + i];
                            },
                            n, cap, lane);
                    }
                }
                nacc += n < cap ? n : cap;
            }
            __syncthreads();
        } else if (cap > 0) {
            nacc = 0;
            for (int t = 0; t < nb; ++t) {
                const int n = bcnt[t];
                const size_t so = (size_t)t * (size_t)A.slot_cap;
                int* oi = A.loc_idx + obase + nacc;
                double* orh = A.loc_rho + obase + nacc;
                if (n <= cap) {
                    for (int i = tid; i < n; i += kSelT) {
                        oi[i] = sob[so + i];
                        orh[i] = srho[so + i];
                    }
                    nacc += n;
                } else {
                    nacc += cap_stream(
                        S, G, [&](int i) { return skey[so + i]; },
                        [&](int slot, int i) {
                            oi[slot] = sob[so + i];
                            orh[slot] = srho[so + i];
                        },
                        n, cap, tid);
                }
            }
        }
        if (tid == 0) {
            A.loc_off[b] = obase;
            A.loc_cnt[b] = nacc;
        }
        __syncthreads();
    }
}

// Nearest-first cap of the global grid's per-point lists: a list longer than
// the cap keeps the cap entries of largest rho (ties: the earlier list entry)
// in the reference's storage-slot order, by the cap of
// scale/letkf/letkf_tools.f90:1041-1156 with the key 2^63 - 1 - bits(rho)
// (rho > 0; a smaller key is a larger rho).  list_cap_wave: one wave per point
// (persistent), caps up to LEKF_CAP_LDS; list_cap: one workgroup per point.
__device__ __forceinline__ unsigned long long rho_key(double rho)
{
    return 0x7fffffffffffffffull - (unsigned long long)__double_as_longlong(rho);
}

__global__ __launch_bounds__(64) void list_cap_wave(ListCapArgs A)
{
    __shared__ WaveCapLds L;
    const int lane = threadIdx.x;
    const long npt = (long)A.ncols * A.nlev;
    for (long pi = blockIdx.x; pi < npt; pi += gridDim.x) {
        const long b = (pi / A.ncols) * (long)A.nij + A.col_begin + pi % A.ncols;
        const long off = A.in_off[b];
        const int n = A.in_cnt[b];
        const long ob = b * (long)A.cap;
        const int* ii = A.in_idx + off;
        const double* ir = A.in_rho + off;
        int kept = n;
        if (n <= A.cap) {
            for (int i = lane; i < n; i += 64) {
                A.out_idx[ob + i] = ii[i];
                A.out_rho[ob + i] = ir[i];
            }
        } else {
            kept = wave_cap(
                L, [&](int i) { return rho_key(ir[i]); },
                [&](int slot, int i) {
                    A.out_idx[ob + slot] = ii[i];
                    A.out_rho[ob + slot] = ir[i];
                },
                n, A.cap, lane);
        }
        if (lane == 0) {
            A.out_off[b] = ob;
            A.out_cnt[b] = kept;
        }
    }
}

__global__ __launch_bounds__(kSelT) void list_cap(ListCapArgs A)
{
    __shared__ SelLds S;
    const int tid = threadIdx.x;
    CapGlobal G;
    G.key = A.g_key ? A.g_key + (size_t)blockIdx.x * 2 * A.cap : nullptr;
    G.pos = A.g_pos ? A.g_pos + (size_t)blockIdx.x * 2 * A.cap : nullptr;
    G.slot = A.g_slot ? A.g_slot + (size_t)blockIdx.x * 2 * A.cap : nullptr;
    const long npt = (long)A.ncols * A.nlev;
    for (long pi = blockIdx.x; pi < npt; pi += gridDim.x) {
        const long b = (pi / A.ncols) * (long)A.nij + A.col_begin + pi % A.ncols;
        const long off = A.in_off[b];
        const int n = A.in_cnt[b];
        const long ob = b * (long)A.cap;
        int kept = n;
        if (n <= A.cap) {
            for (int i = tid; i < n; i += kSelT) {
                A.out_idx[ob + i] = A.in_idx[off + i];
                A.out_rho[ob + i] = A.in_rho[off + i];
            }
        } else {
            const int* ii = A.in_idx + off;
            const double* ir = A.in_rho + off;
            kept = cap_stream(
                S, G, [&](int i) { return rho_key(ir[i]); },
                [&](int slot, int i) {
                    A.out_idx[ob + slot] = ii[i];
                    A.out_rho[ob + slot] = ir[i];
                },
                n, A.cap, tid);
        }
        if (tid == 0) {
            A.out_off[b] = ob;
            A.out_cnt[b] = kept;
        }
        __syncthreads();
    }
}

// per point: candidate count of the rectangle -> list capacity
__global__ void reg_capacity(RegionalArgs A, int* capn, int* maxcand)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long npt = (long)A.nx * A.ny * A.nlev;
    if (b >= npt) return;
    const int nij = A.nx * A.ny;
    const int ij = (int)(b % nij);
    const double ri = (double)(ij % A.nx + 1), rj = (double)(ij / A.nx + 1);
    int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
    int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
    if (imin < 1) imin = 1;
    if (imax > A.nx) imax = A.nx;
    if (jmin < 1) jmin = 1;
    if (jmax > A.ny) jmax = A.ny;
    long n = 0;
    for (int j = jmin; j <= jmax && imin <= imax; ++j) {
        const long rowbase = (long)(j - 1) * A.nx;
        const int beg = (imin > 1) ? A.cellcum[rowbase + imin - 2] : (rowbase > 0 ? A.cellcum[rowbase - 1] : 0);
        n += A.cellcum[rowbase + imax - 1] - beg;
    }
    atomicMax(maxcand, (int)n);
    if (A.nobs_max > 0) {
        const long c = (long)A.nobs_max * A.nbkt;
        n = n < c ? n : c;
    }
    capn[b] = (int)n;
}

// obs keys: cell (ceiling(ri-0.5), ceiling(rj-0.5)) clamped; qc-rejected obs sort last
__global__ void reg_keys(RegObsArgs A)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= A.nobs) return;
    unsigned long long key = ~0ull;
    if (!A.qc || A.qc[n] == A.qc_keep) {
        int i = (int)ceil(A.ri[n] - 0.5), j = (int)ceil(A.rj[n] - 0.5);
        i = i < 1 ? 1 : (i > A.nx ? A.nx : i);
        j = j < 1 ? 1 : (j > A.ny ? A.ny : j);
        const unsigned cell = (unsigned)((i - 1) + A.nx * (j - 1));
        key = ((unsigned long long)cell << 32) | (unsigned)n;
        atomicAdd(&A.cell_cnt[cell], 1);
    }
    A.keys[n] = key;
}

__global__ void reg_scatter(RegObsArgs A, int nbinned)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nbinned) return;
    const int n = (int)(A.keys[m] & 0xffffffffull);
    A.perm[m] = n;
    A.o_ri[m] = A.ri[n];
    A.o_rj[m] = A.rj[n];
    A.o_z[m] = A.z[n];
    A.o_err[m] = A.err[n];
    A.o_rdiag[m] = A.err[n] * A.err[n];
    A.o_dep[m] = A.dep[n];
    A.o_bkt[m] = A.bkt[n];
    for (int i = 0; i < A.ne; ++i) {
        const double v = A.hdxf[(size_t)n * A.h_so + (size_t)i * A.h_sm];
        A.o_hdxf[(size_t)m * A.ne + i] = v;
        if (A.o_hdxf32) A.o_hdxf32[(size_t)m * A.ne + i] = (float)v;
    }
}

__global__ __launch_bounds__(1024) void scan_inclusive_int(const int* in, int* out, int n)
{
    __shared__ int s[1024];
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x];
        const int tot = s[1023];
        __syncthreads();
        carry += tot;
    }
}

__global__ __launch_bounds__(1024) void scan_exclusive_long(const int* in, long* out, long n)
{
    __shared__ long s[1024];
    long carry = 0;
    for (long base = 0; base < n; base += 1024) {
        const long i = base + threadIdx.x;
        const long v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const long t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x] - v;
        const long tot = s[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

}  // namespace

hipError_t lekf_reg_keys_launch(const RegObsArgs& A, hipStream_t st)
{
    if (A.nobs > 0)
        hipLaunchKernelGGL(reg_keys, dim3((A.nobs + 255) / 256), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_reg_scatter_launch(const RegObsArgs& A, int nbinned, int ncell, int* cellcum,
                                   hipStream_t st)
{
    if (nbinned > 0)
        hipLaunchKernelGGL(reg_scatter, dim3((nbinned + 255) / 256), dim3(256), 0, st, A, nbinned);
    hipLaunchKernelGGL(scan_inclusive_int, dim3(1), dim3(1024), 0, st, A.cell_cnt, cellcum, ncell);
    return hipGetLastError();
}

hipError_t lekf_reg_capacity_launch(const RegionalArgs& A, int* capn, long* cap_off, int* maxcand,
                                    hipStream_t st)
{
    const long npt = (long)A.nx * A.ny * A.nlev;
    hipLaunchKernelGGL(reg_capacity, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st, A, capn,
                       maxcand);
    hipLaunchKernelGGL(scan_exclusive_long, dim3(1), dim3(1024), 0, st, capn, cap_off, npt);
    return hipGetLastError();
}

// workgroup slots of the persistent selection (LDS allows up to four per CU)
int lekf_reg_select_slots()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int per = g_lekf_tune.sel_per_cu < 1 ? 1 : (g_lekf_tune.sel_per_cu > 4 ? 4 : g_lekf_tune.sel_per_cu);
    return per * (ncu > 0 ? ncu : 256);
}

int lekf_reg_select_slots_max()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    return 4 * (ncu > 0 ? ncu : 256);
}

hipError_t lekf_list_cap_launch(const ListCapArgs& A, int nslots, hipStream_t st)
{
    const long npt = (long)A.ncols * A.nlev;
    if (npt <= 0) return hipSuccess;
    KTime kt("list_cap", st);
    if (A.cap <= LEKF_CAP_LDS) {
        // one wave per point: as many as the LDS lets a CU hold
        int dev = 0, ncu = 256, per_cu = 1;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, list_cap_wave, 64, 0);
        long g = (long)(per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 256);
        if (g > npt) g = npt;
        hipLaunchKernelGGL(list_cap_wave, dim3((unsigned)g), dim3(64), 0, st, A);
        return hipGetLastError();
    }
    const long g = npt < nslots ? npt : nslots;
    hipLaunchKernelGGL(list_cap, dim3((unsigned)g), dim3(kSelT), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_reg_select_launch(const RegionalArgs& A, long npt, int nslots, hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    const long g = npt < nslots ? npt : nslots;
    KTime kt("select", st);
    hipLaunchKernelGGL(reg_select, dim3((unsigned)g), dim3(kSelT), 0, st, A, npt);
    return hipGetLastError();
}

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

struct SelLds {
    CapLds cap;
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
// The cap by one wave (caps up to LEKF_CAP_LDS), fed chunk by chunk with the
// candidates of one bucket stream in scan order.
//
// The kept set S (the cap smallest (key, scan position) of the stream so
// far) is held ascending in key[], with each entry's storage slot in slot[];
// so[s] (caller's memory) is the candidate in storage slot s (an obs or list
// position, which increases in scan order).  The first cap candidates fill
// slots 0..cap-1; each 64 of them are sorted across the lanes and merged
// into the sorted prefix as they arrive.  A later candidate can enter only if
// its key is below the key tau of S's last entry (an earlier candidate with
// an equal key precedes it), and tau only falls, so candidates are filtered
// against tau into a queue and each block of 64 queued candidates is merged
// at once:
// * a candidate enters iff fewer than cap of S and of the earlier queued
//   candidates precede it; all of S precedes it in scan order, so its count
//   in S is the number of keys <= its key, and the earlier queued candidates
//   before it come from a 64-lane sort by (key, lane);
// * the entries of S move up by the entrants before them, and the top e of
//   S and the entrants leave;
// * the i-th entrant in scan order takes the slot of the i-th largest entry
//   that left: the reference evicts the current last entry each time, and
//   the last key only falls; an entrant that leaves in its own block hands
//   its slot on (pointer jumping).
// ---------------------------------------------------------------------------

// order of the lanes' LDS traffic within the wave (no cross-wave sharing)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct WaveCap {
    unsigned long long key[LEKF_CAP_LDS];
    unsigned short slot[LEKF_CAP_LDS];
    unsigned long long qkey[128];  // queued candidates, scan order
    int qo[128];
    int ev[64];                    // entries that left a block, ascending: slot, or -(entrant + 1)
    int slotn[64];
    int link[64];
};

struct WcState {
    int n;                   // stream candidates so far
    int m;                   // sorted prefix of S while it fills
    int q;                   // queued candidates
    unsigned long long tau;  // key of S's last entry (once n >= cap)
};

// ascending bitonic sort of the 64 lanes' (key, id) pairs (ids distinct)
__device__ __forceinline__ void wave_sort64(unsigned long long& k, int& id, int lane)
{
#pragma unroll
    for (int sz = 2; sz <= 64; sz <<= 1) {
#pragma unroll
        for (int j = sz >> 1; j > 0; j >>= 1) {
            const unsigned lo = (unsigned)lane_xor((int)(unsigned)k, j, lane);
            const unsigned hi = (unsigned)lane_xor((int)(unsigned)(k >> 32), j, lane);
            const unsigned long long ok = ((unsigned long long)hi << 32) | lo;
            const int oid = lane_xor(id, j, lane);
            const bool up = (lane & sz) == 0;
            const bool lower = (lane & j) == 0;
            const bool oless = ok < k || (ok == k && oid < id);
            if (lower == up ? oless : !oless) {
                k = ok;
                id = oid;
            }
        }
    }
}

// number of the first n keys (ascending) that are <= y
__device__ __forceinline__ int wc_count_le(const WaveCap& L, int n, unsigned long long y)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L.key[mid] <= y)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Entries 0..n-1 of key[]/slot[] move up by the new entries placed before
// them (lanes with ent, at count c1 of the old entries: before old entry i
// iff c1 <= i); an entry that lands at lim or above goes to ev[] instead
// (its slot).  Windows of 64 go top down, so a window's writes only reach
// entries that already moved, and each lane holds one entry at a time.
__device__ __forceinline__ void wc_shift(WaveCap& L, int n, int lim, bool ent, int c1, int lane)
{
    const unsigned long long em = __ballot(ent);
    if (!em || n <= 0) return;
    const int c1min = __builtin_amdgcn_readlane(c1, __builtin_ctzll(em));
    for (int lo = ((n - 1) >> 6) << 6; lo >= 0 && lo + 64 > c1min; lo -= 64) {
        const int i = lo + lane;
        unsigned long long k = 0;
        int s = 0;
        if (i < n) {
            k = L.key[i];
            s = L.slot[i];
        }
        int ptr = __popcll(__ballot(ent && c1 < lo));
        unsigned long long wm = __ballot(ent && c1 >= lo && c1 < lo + 64);
        while (wm) {
            const int bl = __builtin_ctzll(wm);
            wm &= wm - 1;
            const int d = __builtin_amdgcn_readlane(c1, bl) - lo;
            ptr += lane >= d ? 1 : 0;
        }
        if (i < n && ptr > 0) {
            const int ni = i + ptr;
            if (ni < lim) {
                L.key[ni] = k;
                L.slot[ni] = (unsigned short)s;
            } else {
                L.ev[ni - lim] = s;
            }
        }
    }
}

// While S fills: the nt (<= 64) entries at m.. (slots m.., unsorted) join
// the sorted prefix 0..m-1.
__device__ void wc_fill_merge(WaveCap& L, int m, int nt, int lane)
{
    unsigned long long sk = lane < nt ? L.key[m + lane] : ~0ull;
    int sid = lane;
    wave_sort64(sk, sid, lane);  // the nt entries take lanes 0..nt-1 (ties: slot order)
    const bool v = lane < nt;
    const int c1 = v ? wc_count_le(L, m, sk) : m;
    wave_sync();
    wc_shift(L, m, m + nt, v, c1, lane);
    wave_sync();
    if (v) {
        L.key[c1 + lane] = sk;
        L.slot[c1 + lane] = (unsigned short)(m + sid);
    }
    wave_sync();
}

// One block: queued candidates 0..np-1 (np <= 64) against the full set S.
__device__ void wc_block(WaveCap& L, int* so, int np, int cap, int lane, unsigned long long& tau)
{
    const bool valid = lane < np;
    const unsigned long long y = valid ? L.qkey[lane] : ~0ull;
    const bool live = valid && y < tau;  // tau has fallen below it since it was queued: out
    const int c1 = live ? wc_count_le(L, cap, y) : cap;
    // sort the live candidates by (key, lane): lane L holds the L-th
    unsigned long long sk = live ? y : ~0ull;
    int sid = lane;
    wave_sort64(sk, sid, lane);
    const bool slive = sk != ~0ull;
    unsigned long long pm = slive ? (1ull << sid) : 0ull;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long t = __shfl_up(pm, d, 64);
        if (lane >= d) pm |= t;
    }
    // earlier (in scan order) live candidates sorted before this one
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
    wave_sync();
    wc_shift(L, cap, cap, ents, c1s, lane);
    wave_sync();
    if (ents) {
        if (U < cap)
            L.key[U] = sk;
        else
            L.ev[U - cap] = -(t + 1);
    }
    wave_sync();
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
    if (ents && U < cap && sl >= 0) {
        L.slot[U] = (unsigned short)sl;
        so[sl] = L.qo[sid];
    }
    wave_sync();
    tau = L.key[cap - 1];
}

// One chunk of the stream: lane l's position is a stream candidate iff ok
// (key k, candidate o); positions are in scan order across the lanes.
__device__ void wc_push(WaveCap& L, int* so, WcState& S, bool ok, unsigned long long k, int o,
                        int cap, int lane)
{
    const unsigned long long am = __ballot(ok);
    if (!am) return;
    const int rank = __popcll(am & lane_lt(lane));
    const int n0 = S.n;
    S.n += __popcll(am);
    bool pend = ok;
    if (n0 < cap) {
        if (ok && n0 + rank < cap) {
            const int f = n0 + rank;
            L.key[f] = k;
            so[f] = o;
            pend = false;
        }
        wave_sync();
        const int nf = S.n < cap ? S.n : cap;
        while (nf - S.m >= 64 || (S.n >= cap && S.m < cap)) {
            const int nt = nf - S.m < 64 ? nf - S.m : 64;
            wc_fill_merge(L, S.m, nt, lane);
            S.m += nt;
        }
        if (S.n < cap) return;
        S.tau = L.key[cap - 1];
        if (S.n == cap) return;
    }
    const bool pot = pend && k < S.tau;
    const unsigned long long pm = __ballot(pot);
    if (!pm) return;
    if (pot) {
        const int qp = S.q + __popcll(pm & lane_lt(lane));
        L.qkey[qp] = k;
        L.qo[qp] = o;
    }
    S.q += __popcll(pm);
    wave_sync();
    if (S.q >= 64) {
        wc_block(L, so, 64, cap, lane, S.tau);
        // the rest of the queue moves to its front
        const bool mv = 64 + lane < S.q;
        unsigned long long k2 = 0;
        int o2 = 0;
        if (mv) {
            k2 = L.qkey[64 + lane];
            o2 = L.qo[64 + lane];
        }
        wave_sync();
        if (mv) {
            L.qkey[lane] = k2;
            L.qo[lane] = o2;
        }
        S.q -= 64;
        wave_sync();
    }
}

// End of the stream: the last queued block.  Returns the kept count; the
// kept candidates are so[0 .. kept) in slot order (visible to the wave's
// loads after a workgroup-scope fence).
__device__ int wc_finish(WaveCap& L, int* so, WcState& S, int cap, int lane)
{
    if (S.q > 0) wc_block(L, so, S.q, cap, lane, S.tau);
    S.q = 0;
    wave_sync();
    return S.n < cap ? S.n : cap;
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
        if (cap > 0) {
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

// Capped selection with the kept sets in LDS (0 < nobs_max <= LEKF_CAP_LDS):
// persistent, one point per workgroup, wave t caps bucket t's stream (NW =
// nbkt waves).  A bucket's candidates in the point's rectangle are, row by
// row, one contiguous range of the bucket-by-bucket obs copy, in the
// reference's scan order restricted to the bucket; each is evaluated once
// (eval_rec) and fed to the wave's cap.  The lists are written bucket by
// bucket in slot order, rho recomputed from the kept obs.
struct RowTab {
    int beg[64];
    int pre[65];
};

constexpr int kTileX = 16, kTileY = 8;  // columns of a selection tile

template <int NW>
__global__ __launch_bounds__(64 * NW) void reg_select_wave(RegionalArgs A, long npt)
{
    __shared__ WaveCap L[NW];
    __shared__ RowTab R[NW];
    __shared__ int kept_s[NW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nij = A.nx * A.ny;
    const int cap = A.nobs_max;
    const int* cc = A.bcc + (size_t)wv * nij;
    int bs = 0;  // bucket wv's first position in the bucket-by-bucket copy
    for (int u = 0; u < wv; ++u) bs += A.bcc[(size_t)u * nij + nij - 1];
    WaveCap& W = L[wv];
    RowTab& T = R[wv];
    int* so = A.wc_so + ((size_t)blockIdx.x * NW + wv) * LEKF_CAP_LDS;  // slot -> obs
    const double r2cut = kNdMax * A.sigma_h * A.sigma_h * (1.0 + 1e-9);
    // the bucket's smallest err^2 (NaN for an empty bucket: never used)
    const double e2min = A.err2_min ? __longlong_as_double((long long)A.err2_min[wv]) : 0.0;
    const float ish2 = (float)(1.0 / (A.sigma_h * A.sigma_h));
    const float isv2 = (float)(1.0 / (A.sigma_v * A.sigma_v));

    // XCD-aware tile order: workgroup w runs on XCD w % 8 (round-robin
    // dispatch); the XCD's workgroups take the points of one kTileX x kTileY
    // block of columns at one level together, so the obs of the block's
    // rectangles stay in that XCD's L2.  Group x (x < ng = min(8, grid)) takes
    // the tiles t = x (mod ng), its workgroups the tile's points in turn.
    const int ng = gridDim.x < 8 ? (int)gridDim.x : 8;
    const int xcd = (int)blockIdx.x % ng, wj = (int)blockIdx.x / ng;
    const int nwg = ((int)gridDim.x - xcd + ng - 1) / ng;
    const int tnx = (A.nx + kTileX - 1) / kTileX, tny = (A.ny + kTileY - 1) / kTileY;
    const long tpl = (long)tnx * tny;  // tiles per level
    const long bend = A.pt_begin + npt;
    const long t_first = (A.pt_begin / nij) * tpl, t_last = ((bend - 1) / nij + 1) * tpl - 1;
    const long t0 = t_first + ((xcd - t_first % ng) + ng) % ng;
    for (long q = wj;; q += nwg) {
        const long t = t0 + (q / (kTileX * kTileY)) * ng;
        if (t > t_last) break;
        const int w = (int)(q % (kTileX * kTileY));
        const int lev = (int)(t / tpl), tl = (int)(t % tpl);
        const int ix = (tl % tnx) * kTileX + w % kTileX, iy = (tl / tnx) * kTileY + w / kTileX;
        if (ix >= A.nx || iy >= A.ny) continue;
        const int ij = iy * A.nx + ix;
        const long b = (long)lev * nij + ij;  // global point
        if (b < A.pt_begin || b >= bend) continue;
        const double ri = (double)(ix + 1), rj = (double)(iy + 1), rz = A.zlev[lev];
        int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
        int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
        if (imin < 1) imin = 1;
        if (imax > A.nx) imax = A.nx;
        if (jmin < 1) jmin = 1;
        if (jmax > A.ny) jmax = A.ny;
        WcState S{0, 0, 0, ~0ull};
        if (imin <= imax) {
            for (int j0 = jmin; j0 <= jmax; j0 += 64) {
                const int nr = jmax - j0 + 1 < 64 ? jmax - j0 + 1 : 64;
                int beg = 0, len = 0;
                if (lane < nr) {
                    // the row's cells that can hold an obs within the cutoff: a
                    // cell's obs lie on its side of the faces at i +- 0.5, so one
                    // whose nearest face is beyond the cutoff radius holds none
                    // (nd >= nd_h^2 > kNdMax); a 1e-9 margin covers rounding
                    const int j = j0 + lane;
                    const double dyg = fmax(0.0, fabs((double)j - rj) - 0.5) * A.dy;
                    const double rem = r2cut - dyg * dyg;
                    int ilo = imin, ihi = imax;
                    if (rem < 0.0) {
                        ihi = ilo - 1;
                    } else {
                        const double rx = sqrt(rem) / A.dx + 0.5;
                        const int a = (int)ceil(ri - rx), z = (int)floor(ri + rx);
                        ilo = a > imin ? a : imin;
                        ihi = z < imax ? z : imax;
                    }
                    if (ilo <= ihi) {
                        const int rowbase = (j - 1) * A.nx;
                        const int lo = rowbase + ilo - 1, hi = rowbase + ihi - 1;
                        const int c0 = lo > 0 ? cc[lo - 1] : 0;
                        beg = bs + c0;
                        len = cc[hi] - c0;
                    }
                }
                int pre = len;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int v = __shfl_up(pre, d, 64);
                    if (lane >= d) pre += v;
                }
                T.beg[lane] = beg;
                T.pre[lane + 1] = pre;
                if (lane == 0) T.pre[0] = 0;
                const int tot = __shfl(pre, 63, 64);
                wave_sync();
                // candidate q of the row group -> its position; loads of the next
                // chunk are issued before the current one is evaluated
                int r0 = 0;
                ObsRec nrec;
                int no = -1;
                auto fetch = [&](int qb) {
                    while (r0 < nr && T.pre[r0 + 1] <= qb) ++r0;
                    const int q = qb + lane;
                    no = -1;
                    if (q < tot) {
                        int r = r0;
                        while (T.pre[r + 1] <= q) ++r;
                        const int p = T.beg[r] + (q - T.pre[r]);
                        const double2 rij = *reinterpret_cast<const double2*>(A.brec + 4 * (size_t)p);
                        const double2 zer = *reinterpret_cast<const double2*>(A.brec + 4 * (size_t)p + 2);
                        nrec.ri = rij.x;
                        nrec.rj = rij.y;
                        nrec.z = zer.x;
                        nrec.err = zer.y;
                        nrec.b = wv;
                        no = A.bo[p];
                    }
                };
                fetch(0);
                for (int q0 = 0; q0 < tot; q0 += 64) {
                    if (S.n >= cap && (double)(j0 + r0) - rj > 0.5) {
                        // rows from this chunk's first one on lie beyond the
                        // point: an obs of row j has |rj_o - rj| > j - rj - 0.5,
                        // so nd > ((j - rj - 0.5) dy / sigma_h)^2 and its key
                        // err^2 / rho exceeds err2_min exp(nd / 2); once that
                        // bound (less a 1e-9 margin for rounding) reaches tau,
                        // which only falls, no later candidate can enter
                        const double g = ((double)(j0 + r0) - rj - 0.5) * A.dy / A.sigma_h;
                        const double lb = e2min * exp(0.5 * g * g) * (1.0 - 1e-9);
                        if (lb >= __longlong_as_double((long long)S.tau)) break;
                    }
                    const ObsRec cur = nrec;
                    const int co = no;
                    if (q0 + 64 < tot)
                        fetch(q0 + 64);
                    else
                        no = -1;
                    Cand c;
                    c.ok = false;
                    c.bits = 0;
                    // once the kept set is full, a float bound in the log2 domain
                    // (errors near 1e-5 for candidates within the cutoff, against
                    // a 1e-3 margin) skips the exact evaluation of a candidate whose
                    // key certainly exceeds tau: it could not enter, and past the
                    // fill the stream count no longer matters
                    bool need = co >= 0;
                    if (need && S.n >= cap) {
                        const float fx = (float)((ri - cur.ri) * A.dx);
                        const float fy = (float)((rj - cur.rj) * A.dy);
                        const float fz = (float)(cur.z - rz);
                        const float nd = (fx * fx + fy * fy) * ish2 + fz * fz * isv2;
                        const float l2 = __log2f((float)(cur.err * cur.err)) + 0.72134752f * nd;
                        need = !(l2 > __log2f((float)__longlong_as_double((long long)S.tau)) + 1e-3f);
                    }
                    if (need) c = eval_rec(A, cur, co, ri, rj, rz);
                    wc_push(W, so, S, c.ok, c.bits, co, cap, lane);
                }
                wave_sync();
            }
        }
        const int kept = wc_finish(W, so, S, cap, lane);
        if (lane == 0) kept_s[wv] = kept;
        __syncthreads();
        int off = 0, tot = 0;
        for (int u = 0; u < NW; ++u) {
            const int k = kept_s[u];
            off += u < wv ? k : 0;
            tot += k;
        }
        const long obase = A.list_off[b] - A.list_base;  // relative to this chunk's list buffer
        for (int sl = lane; sl < kept; sl += 64) {
            const int o = so[sl];
            const Cand c = eval_rec(A, load_obs(A, o), o, ri, rj, rz);
            A.loc_idx[obase + off + sl] = o;
            A.loc_rho[obase + off + sl] = c.rho;
        }
        if (threadIdx.x == 0) {
            A.loc_off[b] = obase;
            A.loc_cnt[b] = tot;
        }
        __syncthreads();
    }
}

// The bucket-by-bucket copy of the sorted obs: per cell, its obs of each
// bucket (count), then each bucket's inclusive scan over cells, then the copy
// (every cell's obs in order, each to its bucket's next position).
__global__ void reg_bucket_count(const int* cellcum, const int* obkt, int ncell, int nbkt,
                                 int* bcnt)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int n[LEKF_REG_NB] = {0, 0, 0, 0};
    const int beg = c > 0 ? cellcum[c - 1] : 0, end = cellcum[c];
    for (int i = beg; i < end; ++i) {
        const int t = obkt[i];
        if (t >= 0 && t < LEKF_REG_NB) ++n[t];
    }
    for (int t = 0; t < nbkt; ++t) bcnt[(size_t)t * ncell + c] = n[t];
}

__global__ void reg_bucket_scatter(const int* cellcum, const double* ori, const double* orj,
                                   const double* oz, const double* oerr, const int* obkt,
                                   int ncell, int nbkt, const int* bcc, double* brec, int* bo,
                                   unsigned long long* err2_min)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int pos[LEKF_REG_NB] = {0, 0, 0, 0};
    int bs = 0;
    for (int t = 0; t < nbkt; ++t) {
        const int* cc = bcc + (size_t)t * ncell;
        pos[t] = bs + (c > 0 ? cc[c - 1] : 0);
        bs += cc[ncell - 1];
    }
    const int beg = c > 0 ? cellcum[c - 1] : 0, end = cellcum[c];
    unsigned long long e2[LEKF_REG_NB] = {~0ull, ~0ull, ~0ull, ~0ull};
    for (int i = beg; i < end; ++i) {
        const int t = obkt[i];
        if (t < 0 || t >= nbkt) continue;
        const int p = pos[t]++;
        double* r = brec + 4 * (size_t)p;
        r[0] = ori[i];
        r[1] = orj[i];
        r[2] = oz[i];
        r[3] = oerr[i];
        bo[p] = i;
        // err^2 >= 0: its bits order as the values do
        const unsigned long long b2 = (unsigned long long)__double_as_longlong(rn_mul(oerr[i], oerr[i]));
        for (int u = 0; u < LEKF_REG_NB; ++u)
            if (u == t && b2 < e2[u]) e2[u] = b2;
    }
    for (int u = 0; u < nbkt && u < LEKF_REG_NB; ++u)
        if (e2[u] != ~0ull) atomicMin(&err2_min[u], e2[u]);
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
    __shared__ WaveCap L;
    __shared__ int so[LEKF_CAP_LDS];
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
            WcState S{0, 0, 0, ~0ull};
            double rn = lane < n ? ir[lane] : 0.0;
            for (int i0 = 0; i0 < n; i0 += 64) {
                const double rc = rn;
                const int i = i0 + lane;
                rn = i + 64 < n ? ir[i + 64] : 0.0;
                wc_push(L, so, S, i < n, rho_key(rc), i, A.cap, lane);
            }
            kept = wc_finish(L, so, S, A.cap, lane);
            for (int sl = lane; sl < kept; sl += 64) {
                const int i = so[sl];
                A.out_idx[ob + sl] = ii[i];
                A.out_rho[ob + sl] = ir[i];
            }
            wave_sync();
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

// a CU holds at most 16 waves of the fused selection (its LDS and registers)
long lekf_reg_wave_slots_max()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    return 16L * (ncu > 0 ? ncu : 256);
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

template <int NW>
static hipError_t launch_select_wave(const RegionalArgs& A, long npt, hipStream_t st)
{
    // as many workgroups as the LDS lets the CUs hold
    int dev = 0, ncu = 256, per_cu = 1;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reg_select_wave<NW>, 64 * NW, 0);
    long g = (long)(per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 256);
    if (g * NW > A.wc_so_waves) g = A.wc_so_waves / NW;  // one slot map per wave
    if (g > npt) g = npt;
    if (g < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reg_select_wave<NW>, dim3((unsigned)g), dim3(64 * NW), 0, st, A, npt);
    return hipGetLastError();
}

hipError_t lekf_reg_select_launch(const RegionalArgs& A, long npt, int nslots, hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    KTime kt("select", st);
    if (lekf_reg_select_fused(A.nobs_max) && A.bcc) {
        switch (A.nbkt) {
        case 1: return launch_select_wave<1>(A, npt, st);
        case 2: return launch_select_wave<2>(A, npt, st);
        case 3: return launch_select_wave<3>(A, npt, st);
        case 4: return launch_select_wave<4>(A, npt, st);
        default: return hipErrorInvalidValue;
        }
    }
    const long g = npt < nslots ? npt : nslots;
    hipLaunchKernelGGL(reg_select, dim3((unsigned)g), dim3(kSelT), 0, st, A, npt);
    return hipGetLastError();
}

hipError_t lekf_reg_buckets_launch(const RegObsArgs& A, int nbinned, const int* cellcum, int nbkt,
                                   int* bcnt, int* bcc, double* brec, int* bo,
                                   unsigned long long* err2_min, hipStream_t st)
{
    (void)nbinned;
    const int ncell = A.nx * A.ny;
    if (ncell <= 0 || nbkt < 1 || nbkt > LEKF_REG_NB) return hipErrorInvalidValue;
    const int nb = (ncell + 255) / 256;
    hipLaunchKernelGGL(reg_bucket_count, dim3(nb), dim3(256), 0, st, cellcum, A.o_bkt, ncell, nbkt,
                       bcnt);
    for (int t = 0; t < nbkt; ++t)
        hipLaunchKernelGGL(scan_inclusive_int, dim3(1), dim3(1024), 0, st, bcnt + (size_t)t * ncell,
                           bcc + (size_t)t * ncell, ncell);
    hipLaunchKernelGGL(reg_bucket_scatter, dim3(nb), dim3(256), 0, st, cellcum, A.o_ri, A.o_rj,
                       A.o_z, A.o_err, A.o_bkt, ncell, nbkt, bcc, brec, bo, err2_min);
    return hipGetLastError();
}

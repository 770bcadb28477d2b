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
// rdiag = err^2/rloc.  With MAX_NOBS_PER_GRID = nobs_max > 0 every bucket
// (the reference's (element, obs type) pair) keeps its nobs_max smallest
// (rdiag, scan position) pairs -- exactly the set the reference's insertion
// sort keeps (strict < keeps earlier obs on ties).  Instead of inserting
// serially, one wave per point finds each bucket's threshold pair with
// radix-digit histogram passes over the 64-bit key (10-bit digits in LDS),
// resolves the last digit bin by sorting at most LEKF_REG_LIST candidates,
// and emits the kept obs bucket by bucket in scan order.  Without a cap the
// accepted obs are emitted in scan order.
#include "lekf_dev.hpp"
#include "lekf_regional.hpp"

namespace {

constexpr double kNdMax = 13.33333333;  // scale/letkf/letkf_obs.f90:30

__device__ __forceinline__ unsigned long long lane_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

struct RegLds {
    int hist[LEKF_REG_NB][1024];
    int cnt[LEKF_REG_NB];
    int need[LEKF_REG_NB];
    int bincnt[LEKF_REG_NB];
    int mode[LEKF_REG_NB];  // 0 keep all, 1 refine, 2 collect, 3 tie scan, 4 resolved
    int shift[LEKF_REG_NB];
    unsigned long long pref[LEKF_REG_NB];
    unsigned long long thr_bits[LEKF_REG_NB];
    int thr_q[LEKF_REG_NB];
    int lcnt[LEKF_REG_NB];
    unsigned long long lbits[LEKF_REG_NB][LEKF_REG_LIST];
    int lq[LEKF_REG_NB][LEKF_REG_LIST];
    int seg_beg[LEKF_REG_SEG];
    int seg_pre[LEKF_REG_SEG + 1];
    int outc[LEKF_REG_NB];
    int base[LEKF_REG_NB + 1];
    int any;
};

struct Cand {
    bool ok;
    int o, b;
    double rho, rd;
    unsigned long long bits;
};

__device__ __forceinline__ Cand eval_cand(const RegionalArgs& A, int o, double ri, double rj,
                                          double rz)
{
    Cand c;
    const double rdx = __dmul_rn(__dsub_rn(ri, A.ori[o]), A.dx);
    const double rdy = __dmul_rn(__dsub_rn(rj, A.orj[o]), A.dy);
    const double nd_h = __ddiv_rn(__dsqrt_rn(__dadd_rn(__dmul_rn(rdx, rdx), __dmul_rn(rdy, rdy))),
                                  A.sigma_h);
    const double nd_v = __ddiv_rn(fabs(__dsub_rn(A.oz[o], rz)), A.sigma_v);
    const double nd = __dadd_rn(__dmul_rn(nd_h, nd_h), __dmul_rn(nd_v, nd_v));
    c.o = o;
    c.ok = !(nd > kNdMax);
    c.b = A.obkt[o];
    c.rho = 0.0;
    c.rd = 0.0;
    c.bits = 0;
    if (c.ok) {
        c.rho = exp(-0.5 * nd);
        const double e = A.oerr[o];
        c.rd = __ddiv_rn(__dmul_rn(e, e), c.rho);
        c.bits = (unsigned long long)__double_as_longlong(c.rd);
    }
    return c;
}

// candidate q of the point's scan (row segments in LDS)
__device__ __forceinline__ int cand_obs(const RegLds& S, int nseg, int q, int& seg)
{
    while (seg + 1 < nseg && S.seg_pre[seg + 1] <= q) ++seg;
    return S.seg_beg[seg] + (q - S.seg_pre[seg]);
}

// Digits of a positive double's bit pattern (monotone in the value): shifts
// 53, 43, 33, 23, 13, 3 take 10 bits (bit 63 is the zero sign bit), shift 0
// the last 3 bits.
__device__ __forceinline__ int digit_width(int shift) { return shift == 0 ? 3 : 10; }

__device__ __forceinline__ int digit(unsigned long long bits, int shift)
{
    return (int)((bits >> shift) & ((1ull << digit_width(shift)) - 1ull));
}

// the bits at and above shift + width(shift) are already fixed in pref;
// shift < 0 means every bit is fixed
__device__ __forceinline__ bool prefix_match(unsigned long long bits, unsigned long long pref,
                                             int shift)
{
    if (shift < 0) return bits == pref;
    const int fixed_from = shift + digit_width(shift);
    if (fixed_from >= 63) return true;
    return (bits >> fixed_from) == (pref >> fixed_from);
}

__device__ __forceinline__ int next_shift(int shift)
{
    return shift == 3 ? 0 : (shift == 0 ? -1 : shift - 10);
}

__global__ __launch_bounds__(64) void reg_select(RegionalArgs A)
{
    __shared__ RegLds S;
    const long b = A.pt_begin + blockIdx.x;  // global point = lev * nij + ij
    const int lane = threadIdx.x;
    const int nij = A.nx * A.ny;
    const int lev = (int)(b / nij), ij = (int)(b % nij);
    const double ri = (double)(ij % A.nx + 1), rj = (double)(ij / A.nx + 1), rz = A.zlev[lev];
    const int nb = A.nbkt;
    const int cap = A.nobs_max;
    int imin = (int)floor(ri - A.dlon_zero), imax = (int)ceil(ri + A.dlon_zero);
    int jmin = (int)floor(rj - A.dlat_zero), jmax = (int)ceil(rj + A.dlat_zero);
    if (imin < 1) imin = 1;
    if (imax > A.nx) imax = A.nx;
    if (jmin < 1) jmin = 1;
    if (jmax > A.ny) jmax = A.ny;
    int nseg = (imin > imax || jmin > jmax) ? 0 : jmax - jmin + 1;
    if (lane == 0) {
        int acc = 0;
        for (int s = 0; s < nseg; ++s) {
            const int j = jmin + s;  // 1-based row
            // nobsgrd(i, j) = cumulative count through cell (i, j); (0, j) = end of row j-1
            const long rowbase = (long)(j - 1) * A.nx;
            const int beg = (imin > 1) ? A.cellcum[rowbase + imin - 2] : (rowbase > 0 ? A.cellcum[rowbase - 1] : 0);
            const int end = A.cellcum[rowbase + imax - 1];
            S.seg_beg[s] = beg;
            S.seg_pre[s] = acc;
            acc += end - beg;
        }
        S.seg_pre[nseg] = acc;
        S.any = 0;
    }
    for (int t = lane; t < LEKF_REG_NB; t += 64) {
        S.cnt[t] = 0;
        S.lcnt[t] = 0;
        S.outc[t] = 0;
        S.mode[t] = 0;
    }
    for (int t = lane; t < LEKF_REG_NB * 1024; t += 64) (&S.hist[0][0])[t] = 0;
    __syncthreads();
    const int ncand = nseg ? S.seg_pre[nseg] : 0;
    const long obase = A.list_off[b] - A.list_base;  // relative to this chunk's list buffer

    if (cap <= 0) {
        // no cap: accepted candidates in scan order
        int seg = 0, n = 0;
        for (int q0 = 0; q0 < ncand; q0 += 64) {
            const int q = q0 + lane;
            Cand c;
            c.ok = false;
            if (q < ncand) c = eval_cand(A, cand_obs(S, nseg, q, seg), ri, rj, rz);
            const unsigned long long m = __ballot(c.ok);
            if (c.ok) {
                const long w = obase + n + __popcll(m & lane_lt(lane));
                A.loc_idx[w] = c.o;
                A.loc_rho[w] = c.rho;
            }
            n += __popcll(m);
        }
        if (lane == 0) {
            A.loc_off[b] = obase;
            A.loc_cnt[b] = n;
        }
        return;
    }

    // pass 0: counts and top-digit histograms per bucket
    {
        int seg = 0;
        for (int q0 = 0; q0 < ncand; q0 += 64) {
            const int q = q0 + lane;
            if (q < ncand) {
                const Cand c = eval_cand(A, cand_obs(S, nseg, q, seg), ri, rj, rz);
                if (c.ok) {
                    atomicAdd(&S.cnt[c.b], 1);
                    atomicAdd(&S.hist[c.b][digit(c.bits, 53)], 1);
                }
            }
        }
    }
    __syncthreads();
    // choose the boundary digit of every bucket over its cap
    if (lane < nb) {
        const int t = lane;
        if (S.cnt[t] > cap) {
            int need = cap, d = 0, acc = 0;
            for (; d < 1024; ++d) {
                if (acc + S.hist[t][d] >= need) break;
                acc += S.hist[t][d];
            }
            S.need[t] = need - acc;
            S.pref[t] = (unsigned long long)d << 53;
            S.bincnt[t] = S.hist[t][d];
            S.shift[t] = 43;
            S.mode[t] = S.bincnt[t] <= LEKF_REG_LIST ? 2 : 1;
        }
    }
    __syncthreads();
    // refine passes while a boundary bin is too large to sort
    for (int pass = 0; pass < 8; ++pass) {
        if (lane == 0) {
            int a = 0;
            for (int t = 0; t < nb; ++t) a |= (S.mode[t] == 1);
            S.any = a;
        }
        __syncthreads();
        if (!S.any) break;
        for (int t = 0; t < nb; ++t)
            if (S.mode[t] == 1)
                for (int u = lane; u < 1024; u += 64) S.hist[t][u] = 0;
        __syncthreads();
        int seg = 0;
        for (int q0 = 0; q0 < ncand; q0 += 64) {
            const int q = q0 + lane;
            if (q < ncand) {
                const Cand c = eval_cand(A, cand_obs(S, nseg, q, seg), ri, rj, rz);
                if (c.ok && S.mode[c.b] == 1 && prefix_match(c.bits, S.pref[c.b], S.shift[c.b]))
                    atomicAdd(&S.hist[c.b][digit(c.bits, S.shift[c.b])], 1);
            }
        }
        __syncthreads();
        if (lane < nb && S.mode[lane] == 1) {
            const int t = lane;
            const int sh = S.shift[t];
            int need = S.need[t], d = 0, acc = 0;
            for (; d < 1024; ++d) {
                if (acc + S.hist[t][d] >= need) break;
                acc += S.hist[t][d];
            }
            S.need[t] = need - acc;
            S.pref[t] |= (unsigned long long)d << sh;
            S.bincnt[t] = S.hist[t][d];
            const int ns = next_shift(sh);
            S.shift[t] = ns;
            if (S.bincnt[t] <= LEKF_REG_LIST)
                S.mode[t] = 2;
            else if (ns < 0)
                S.mode[t] = 3;  // every key bit fixed: equal keys, break ties by scan order
        }
        __syncthreads();
    }
    // collect the boundary bins (sortable) / resolve exact ties in scan order
    {
        int seg = 0, seen[LEKF_REG_NB];
        for (int t = 0; t < LEKF_REG_NB; ++t) seen[t] = 0;
        for (int q0 = 0; q0 < ncand; q0 += 64) {
            const int q = q0 + lane;
            Cand c;
            c.ok = false;
            if (q < ncand) c = eval_cand(A, cand_obs(S, nseg, q, seg), ri, rj, rz);
            bool coll = c.ok && S.mode[c.b] == 2 && prefix_match(c.bits, S.pref[c.b], S.shift[c.b]);
            if (coll) {
                const int pos = atomicAdd(&S.lcnt[c.b], 1);
                if (pos < LEKF_REG_LIST) {
                    S.lbits[c.b][pos] = c.bits;
                    S.lq[c.b][pos] = q;
                }
            }
            for (int t = 0; t < nb; ++t) {
                if (S.mode[t] != 3) continue;
                const bool tie = c.ok && c.b == t && c.bits == S.pref[t];
                const unsigned long long m = __ballot(tie);
                const int before = seen[t];
                const int rank = before + __popcll(m & lane_lt(lane));
                if (tie && rank == S.need[t] - 1) {
                    S.thr_bits[t] = c.bits;
                    S.thr_q[t] = q;
                }
                seen[t] = before + __popcll(m);
            }
        }
    }
    __syncthreads();
    // sort each collected list by (bits, q) and take the need-th element
    for (int t = 0; t < nb; ++t) {
        if (S.mode[t] != 2) continue;
        const int n = S.lcnt[t] < LEKF_REG_LIST ? S.lcnt[t] : LEKF_REG_LIST;
        // odd-even transposition sort, n <= LEKF_REG_LIST
        for (int ph = 0; ph < n; ++ph) {
            for (int i = 2 * lane + (ph & 1); i + 1 < n; i += 128) {
                const unsigned long long a = S.lbits[t][i], c = S.lbits[t][i + 1];
                const int qa = S.lq[t][i], qc = S.lq[t][i + 1];
                if (a > c || (a == c && qa > qc)) {
                    S.lbits[t][i] = c;
                    S.lbits[t][i + 1] = a;
                    S.lq[t][i] = qc;
                    S.lq[t][i + 1] = qa;
                }
            }
            __syncthreads();
        }
        if (lane == 0) {
            S.thr_bits[t] = S.lbits[t][S.need[t] - 1];
            S.thr_q[t] = S.lq[t][S.need[t] - 1];
        }
        __syncthreads();
    }
    if (lane == 0) {
        int acc = 0;
        for (int t = 0; t < nb; ++t) {
            S.base[t] = acc;
            acc += S.cnt[t] < cap ? S.cnt[t] : cap;
        }
        S.base[nb] = acc;
    }
    __syncthreads();
    // emit bucket by bucket in scan order
    {
        int seg = 0;
        int run[LEKF_REG_NB];
        for (int t = 0; t < LEKF_REG_NB; ++t) run[t] = 0;
        for (int q0 = 0; q0 < ncand; q0 += 64) {
            const int q = q0 + lane;
            Cand c;
            c.ok = false;
            if (q < ncand) c = eval_cand(A, cand_obs(S, nseg, q, seg), ri, rj, rz);
            bool keep = false;
            if (c.ok) {
                const int t = c.b;
                keep = (S.mode[t] == 0) || c.bits < S.thr_bits[t] ||
                       (c.bits == S.thr_bits[t] && q <= S.thr_q[t]);
            }
            for (int t = 0; t < nb; ++t) {
                const bool mine = keep && c.b == t;
                const unsigned long long m = __ballot(mine);
                if (mine) {
                    const long w = obase + S.base[t] + run[t] + __popcll(m & lane_lt(lane));
                    A.loc_idx[w] = c.o;
                    A.loc_rho[w] = c.rho;
                }
                run[t] += __popcll(m);
            }
        }
    }
    if (lane == 0) {
        A.loc_off[b] = obase;
        A.loc_cnt[b] = S.base[nb];
    }
}

// per point: candidate count of the rectangle -> list capacity
__global__ void reg_capacity(RegionalArgs A, int* capn)
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
    if (!A.qc || A.qc[n] == 1) {
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

hipError_t lekf_reg_capacity_launch(const RegionalArgs& A, int* capn, long* cap_off, hipStream_t st)
{
    const long npt = (long)A.nx * A.ny * A.nlev;
    hipLaunchKernelGGL(reg_capacity, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st, A, capn);
    hipLaunchKernelGGL(scan_exclusive_long, dim3(1), dim3(1024), 0, st, capn, cap_off, npt);
    return hipGetLastError();
}

hipError_t lekf_reg_select_launch(const RegionalArgs& A, long npt, hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    hipLaunchKernelGGL(reg_select, dim3((unsigned)npt), dim3(64), 0, st, A);
    return hipGetLastError();
}

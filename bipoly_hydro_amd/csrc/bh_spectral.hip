// bh_spectral.hip -- the mixed solve's spectral step for one azimuthal mode
// per workgroup, fused so that the mode's nri x nzi spectral block stays in
// LDS from the forward sine transform to the inverse (see bh_spectral.h).
//
// The separate-kernel form of the same step (two folded sine-transform
// GEMMs, the M_r GEMM with the filter epilogue, the Thomas kernel) moves the
// block through HBM five times; here it is read once (r, M_r, LU factors)
// and written once (delta, w).
//
// Workgroup: 8 waves.  Wave w owns rows j in [32w, 32w + 32) of the block
// in the forward transform and the M_r product, and columns j of the same
// range in the inverse transforms.  The LDS block holds [j][m'] with a row
// pitch of nzi + 4 floats (16-byte aligned rows for the b128 reads of the
// inverse).  The sine transforms of nzi >= 64 and the M_r product are
// split-bf16 products on v_mfma_f32_16x16x32_bf16 (x = hi + lo, three
// products hi.hi + hi.lo + lo.hi summed in FP32, about 2^-17 relative):
// lane l holds row (or column) l % 16 and the eight k values 8 (l / 16) + e,
// result element v of lane l at D[4(l/16) + v][l%16].  The tables are stored
// split, as [tile][32-wide k group][hi, lo][lane][8]; the data operand is
// split as it is read.  nzi = 32 keeps FP32 MFMA (v_mfma_f32_16x16x4f32, A
// lane l: row l%16, k l/16; four k values 4(l/16) + g of a 16-wide k group
// as one 16-byte load).
#include "bh_mixed.h"
#include "bh_residual.h"
#include "bh_spectral.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace bh {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

constexpr int SPEC_THREADS = 512;
constexpr int SPEC_ROWS = 256;  // largest nri

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// lane l: A row l % 16, B column l % 16, eight k values per lane taken in the
// same order from both (any consistent assignment; here k = 8 (l / 16) + e)
__device__ __forceinline__ f4 mma_bf16(bf8 a, bf8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int NZ>
struct Shape {
    static constexpr int HE = NZ / 2;     // even modes = odd modes
    static constexpr int CT = NZ / 16;    // 16-wide spectral column tiles
    static constexpr int KG = HE / 16;    // 16-wide k groups of one half
    static constexpr int ZT = HE / 16;    // 16-row z tiles of the folded inverse
    static constexpr int P = NZ + 4;      // LDS row pitch (floats)
    static constexpr bool BF = NZ >= 64;  // split-bf16 sine transforms possible
    static constexpr int KB = HE / 32;    // 32-wide k groups of one half (BF)
};

// element e of h and l: x = hi + lo in bf16, each rounded to nearest
__device__ __forceinline__ void split_bf16(float x, bf8 &h, bf8 &l, int e)
{
    const __bf16 hi = (__bf16)x;
    h[e] = hi;
    l[e] = (__bf16)(x - (float)hi);
}
// the three products of a split-bf16 multiply-add, small terms first
__device__ __forceinline__ f4 mma3(const bf8 &ah, const bf8 &al, const bf8 &bh, const bf8 &bl, f4 c)
{
    c = mma_bf16(al, bh, c);
    c = mma_bf16(ah, bl, c);
    return mma_bf16(ah, bh, c);
}

// Global-memory views of the argument pointers: loads through them are
// vector-memory loads with their own wait counter (generic pointers become
// flat loads, whose waits also drain the LDS reads).
#define BH_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const BH_GLOBAL T *gld(const void *p)
{
    return (const BH_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ BH_GLOBAL T *gst(void *p)
{
    return (BH_GLOBAL T *)p;
}

// The refinement residual of plane L (FP64, residual_tile_v's arithmetic)
// into LDS as r[k][j] with pitch rp, tiles of 8 K rows, 256 threads per
// tile: thread tid takes column tid % 256 of the tiles tid / 256, + 2, ...
// A cold start (R0 false) reads the faces of X and rh; a warm one (R0: the
// kernel variant of ra.r0 set) the first residual r0 in place of both.
template <bool R0>
__device__ __forceinline__ void spec_residual(const ResidualArgs &ra, int L, int nzi, int tid, float *rl, int rp)
{
    constexpr int RR = 8;
    const int ntiles = nzi / RR, tx = tid & 255;
    const bool add2 = ra.dlt2 != nullptr;
    for (int kt = tid >> 8; kt < ntiles; kt += SPEC_THREADS / 256) {
        const bool edge = kt == 0 || kt == ntiles - 1;
        if (edge) {
            if (add2) residual_tile_v<RR, true, true, true, false, true, R0>(ra, 0, kt, L, tx, rl, rp);
            else residual_tile_v<RR, true, true, true, false, false, R0>(ra, 0, kt, L, tx, rl, rp);
        } else {
            if (add2) residual_tile_v<RR, true, true, false, false, true, R0>(ra, 0, kt, L, tx, rl, rp);
            else residual_tile_v<RR, true, true, false, false, false, R0>(ra, 0, kt, L, tx, rl, rp);
        }
    }
}

// Tile t takes plane t/2 + (t%2) N/2: planes m and m + N/2 (the cosine and
// sine rows of azimuthal mode m) run side by side on one XCD, and share M_r
// and the pivots, stored per mode number mi = L or L - N/2.
// BFT: split-bf16 sine transforms for delta (and, without MR, the forward one)
// R0: the fused residual (a.fuse_res) of a warm refinement pass, from r0
template <int NZ, bool MR, bool ACC, bool BFT, bool R0 = false>
__device__ __forceinline__ void spectral_tile(const SpecArgs &a, float *buf, int tile)
{
    typedef Shape<NZ> S;
    constexpr int P = S::P;
    const int hn = a.nphi / 2;
    const int L = (tile >> 1) + (tile & 1) * hn, mi = L <= hn ? L : L - hn;
    // numr and the thread index pass through an empty asm so that the
    // compiler treats them as new in every tile of the persistent loop:
    // otherwise it hoists hundreds of tile-invariant address offsets out of
    // the loop and spills
    int numr = a.numr;
    asm volatile("" : "+s"(numr));
    const int nri = numr - 2;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int t = lane & 15, q = lane >> 4;
    const int j0 = wave * 32;
    const bool active = j0 < nri;
    float *pq = buf + SPEC_ROWS * P;  // [2][4][NZ] segment maps of the Thomas sweeps
    float *bl = pq + 8 * NZ;          // [SPEC_ROWS] sub-diagonal a_j = betar(j + 2), a_0 = 0
    float *cl = bl + SPEC_ROWS;       // [SPEC_ROWS] super-diagonal c_j = alphar(j + 2), c_{nri-1} = 0
    // their loads fly during the forward transform; LDS is written after it
    const double blv = tid > 0 && tid < nri ? gld<double>(a.betar)[tid + 2] : 0.0;
    const double clv = tid < nri - 1 ? gld<double>(a.alphar)[tid + 2] : 0.0;

    // ---- 1. RT = r^T S, folded: even columns from u = r_a + r_{n-1-a}, odd
    //         from v = r_a - r_{n-1-a}.  Row t of row block b is j0 + 2t + b,
    //         so one 8-byte load gives a lane both blocks' A values.  The
    //         folded table is staged in LDS (in the block's space, free until
    //         the products are written): read once per workgroup, not once
    //         per wave.  The r planes of k group kg + 1 load while group kg is
    //         multiplied.
    const BH_GLOBAL float *r = gld<float>(a.r) + (long)L * NZ * nri + j0 + 2 * t;
    f4 acc[2][S::CT];
    // The first pass (MR) keeps its forward transform in FP32: it feeds w,
    // whose error no refinement pass removes; a correction's transforms are
    // split-bf16, their error contracted by the next pass (or, in the last,
    // below 1e-5 of a correction of about 1e-7 of x)
    if constexpr (BFT && !MR) {
        // every plane this lane needs, loaded at once: k = 32 kb + 8 q + e
        f2 xr[S::KB][8], yr[S::KB][8];
        if (a.fuse_res) {
            // the plane's residual into LDS (pitch nri + 2: the four 16-lane
            // rows of an 8-byte read fall on distinct banks), then the
            // lane's planes from there; the block's space is free again for
            // the table after the second barrier
            const int rp = nri + 2;
            spec_residual<R0>(a.ra, L, NZ, tid, buf, rp);
            __syncthreads();
            if (active) {
#pragma unroll
                for (int kb = 0; kb < S::KB; ++kb)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int pa = 32 * kb + 8 * q + e;
                        xr[kb][e] = *(const f2 *)(buf + pa * rp + j0 + 2 * t);
                        yr[kb][e] = *(const f2 *)(buf + (NZ - 1 - pa) * rp + j0 + 2 * t);
                    }
            }
            __syncthreads();
        } else if (active) {
#pragma unroll
            for (int kb = 0; kb < S::KB; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int pa = 32 * kb + 8 * q + e;
                    xr[kb][e] = *(const BH_GLOBAL f2 *)(r + (long)pa * nri);
                    yr[kb][e] = *(const BH_GLOBAL f2 *)(r + (long)(NZ - 1 - pa) * nri);
                }
        }
        {
            const BH_GLOBAL f4 *sbg = gld<f4>(a.sfb16);
            f4 *sbl = (f4 *)buf;
            for (int i = tid; i < S::HE * NZ / 4; i += SPEC_THREADS) sbl[i] = sbg[i];
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < S::CT; ++c) acc[b][c] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < S::KB; ++kb) {
                bf8 uh[2], ul[2], vh[2], vl[2];
#pragma unroll
                for (int e = 0; e < 8; ++e)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        split_bf16(xr[kb][e][b] + yr[kb][e][b], uh[b], ul[b], e);
                        split_bf16(xr[kb][e][b] - yr[kb][e][b], vh[b], vl[b], e);
                    }
                const bf8 *sb = (const bf8 *)buf + lane;
#pragma unroll
                for (int c = 0; c < S::CT; ++c) {
                    const bf8 bh = sb[((c * S::KB + kb) * 2) * 64], bl = sb[((c * S::KB + kb) * 2 + 1) * 64];
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[b][c] = c < S::CT / 2 ? mma3(uh[b], ul[b], bh, bl, acc[b][c])
                                                  : mma3(vh[b], vl[b], bh, bl, acc[b][c]);
                }
            }
        }
    } else {
    f2 xr[2][4], yr[2][4];
    auto load = [&](int kg, int slot) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int pa = 16 * kg + 4 * q + g;
            xr[slot][g] = *(const BH_GLOBAL f2 *)(r + (long)pa * nri);
            yr[slot][g] = *(const BH_GLOBAL f2 *)(r + (long)(NZ - 1 - pa) * nri);
        }
    };
    if (active) load(0, 0);
    {
        const BH_GLOBAL f4 *sbg = gld<f4>(a.sfb);
        f4 *sbl = (f4 *)buf;
        for (int i = tid; i < S::HE * NZ / 4; i += SPEC_THREADS) sbl[i] = sbg[i];
    }
    __syncthreads();
    if (active) {
        const f4 *sb = (const f4 *)buf + lane;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < S::CT; ++c) acc[b][c] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kg = 0; kg < S::KG; ++kg) {
            const int cur = kg & 1;
            if (kg + 1 < S::KG) load(kg + 1, cur ^ 1);
            // the loads above stay ahead of this group's products (the
            // scheduler would otherwise sink them next to their use)
            __builtin_amdgcn_sched_barrier(0);
            f2 u[4], v[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                u[g] = xr[cur][g] + yr[cur][g];
                v[g] = xr[cur][g] - yr[cur][g];
            }
#pragma unroll
            for (int c = 0; c < S::CT; ++c) {
                const f4 bc = sb[(c * S::KG + kg) * 64];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int b = 0; b < 2; ++b) acc[b][c] = mma(c < S::CT / 2 ? u[g][b] : v[g][b], bc[g], acc[b][c]);
            }
        }
    }
    }
    __syncthreads();  // every wave is done with the table
    if (active) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < S::CT; ++c)
#pragma unroll
                for (int e = 0; e < 4; ++e) buf[(j0 + 2 * (4 * q + e) + b) * P + 16 * c + t] = acc[b][c][e];
    }
    if (tid < SPEC_ROWS) {
        bl[tid] = (float)blv;
        cl[tid] = (float)clv;
    }
    __syncthreads();

    // ---- 2. W = mu . (M_r RT), C = RT - W in place, as split-bf16 products
    //         (x = hi + lo; hi.hi + hi.lo + lo.hi in FP32, v_mfma_f32_16x16x32_bf16):
    //         M_r comes split from the table, RT is split as it is read.  Wave
    //         (rg, cg) = (wave / 2, wave % 2) owns rows 64 rg + [0, 64) and columns
    //         (nzi/2) cg + [0, nzi/2); W stays in registers for its inverse.
    constexpr int CTW = S::CT / 2;  // 16-wide column tiles per wave
    const int rg = wave >> 1, cg = wave & 1;
    const bool fil = rg < nri / 64;
    f4 wacc[4][CTW];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < CTW; ++c) wacc[i][c] = (f4){0.f, 0.f, 0.f, 0.f};
    if (MR) {
        if (fil) {
            const int nkc = nri / 32;
            // fragment (row tile 4 rg + i, chunk kc, part): 64 lanes x 8 bf16
            const BH_GLOBAL bf8 *mrb = gld<bf8>(a.mr) + (long)mi * nri * nri / 4 + (long)4 * rg * nkc * 128 + lane;
            bf8 ah[4], al_[4], nh[4], nl[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ah[i] = mrb[(i * nkc) * 128];
                al_[i] = mrb[(i * nkc) * 128 + 64];
            }
            const float *bcol = buf + 8 * q * P + (NZ / 2) * cg + t;
            for (int kc = 0; kc < nkc; ++kc) {
                if (kc + 1 < nkc) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        nh[i] = mrb[(i * nkc + kc + 1) * 128];
                        nl[i] = mrb[(i * nkc + kc + 1) * 128 + 64];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int c = 0; c < CTW; ++c) {
                    // B[k = 32 kc + 8 q + e][col]: rows of the LDS block, split on the fly
                    bf8 bh, bo;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float x = bcol[(32 * kc + e) * P + 16 * c];
                        const __bf16 h = (__bf16)x;
                        bh[e] = h;
                        bo[e] = (__bf16)(x - (float)h);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        wacc[i][c] = mma_bf16(al_[i], bh, wacc[i][c]);
                        wacc[i][c] = mma_bf16(ah[i], bo, wacc[i][c]);
                        wacc[i][c] = mma_bf16(ah[i], bh, wacc[i][c]);
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    ah[i] = nh[i];
                    al_[i] = nl[i];
                }
            }
        }
        __syncthreads();
        if (fil) {
            const BH_GLOBAL float *mu = gld<float>(a.mu);
#pragma unroll
            for (int c = 0; c < CTW; ++c) {
                const int col = (NZ / 2) * cg + 16 * c + t;
                const float m = mu[col];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int idx = (64 * rg + 16 * i + 4 * q + e) * P + col;
                        const float wv = m * wacc[i][c][e];
                        wacc[i][c][e] = wv;
                        buf[idx] = buf[idx] - wv;
                    }
            }
        }
        __syncthreads();
    }

    // ---- 3. (B_r(L) + lambda_m') D = C along j for every column m', as the
    //         separate Thomas kernel's recurrences (forward y_j = (c_j -
    //         a_j y_{j-1}) ib_j, back x_j = y_j - g_{j+1} x_{j+1}), split into
    //         nri/64 segments of 64 rows: a thread holds its segment of one
    //         column in registers, composes the segment's affine map, the maps
    //         are chained through LDS, and the segment is swept again from its
    //         true entry value.  The back sweep's g_{j+1} = alphar(j + 2) ib_j
    //         comes from the pivots already in registers, so only ib is read,
    //         all 64 loads at once.
    {
        const int nseg = nri / 64;
        const bool tri = tid < nseg * NZ;  // whole waves: nzi is a multiple of 32
        const int m = tid % NZ, sg = tid / NZ, jlo = sg * 64;
        const long fb = (long)mi * nri * NZ + (long)jlo * NZ + m;
        float *col = buf + jlo * P + m;
        const float *al = bl + jlo;
        float fv[64], cv[64];
        if (tri) {
            const BH_GLOBAL float *ib = gld<float>(a.ibet) + fb;
#pragma unroll
            for (int u = 0; u < 64; ++u) fv[u] = ib[u * NZ];
#pragma unroll
            for (int u = 0; u < 64; ++u) cv[u] = col[u * P];
            __builtin_amdgcn_sched_barrier(0);  // all loads in flight before the first use
            float Pf = 0.f, Qf = 1.f;  // y_{jlo + 63} = Pf + Qf y_{jlo - 1}
#pragma unroll
            for (int u = 0; u < 64; ++u) {
                const float aj = al[u];
                Pf = (cv[u] - aj * Pf) * fv[u];
                Qf = -aj * Qf * fv[u];
            }
            pq[sg * NZ + m] = Pf;
            pq[(4 + sg) * NZ + m] = Qf;
        }
        __syncthreads();
        float Pb = 0.f, Qb = 1.f;  // x_{jlo} = Pb + Qb x_{jlo + 64}
        if (tri) {
            float y = 0.f;  // y_{jlo - 1} (a_0 = 0 in the first segment)
            for (int k = 0; k < sg; ++k) y = pq[k * NZ + m] + pq[(4 + k) * NZ + m] * y;
#pragma unroll
            for (int u = 0; u < 64; ++u) {
                y = (cv[u] - al[u] * y) * fv[u];
                cv[u] = y;
            }
            // back over the same rows, high j to low: fv[u] = g_{j+1} for
            // j = jlo + 63 - u (zero past the last row)
            const float *cu = cl + jlo;
#pragma unroll
            for (int u = 0; u < 32; ++u) {  // reverse in place: fv[u] <- c_{63-u} ib_{63-u}
                const float lo = fv[u], hi = fv[63 - u];
                fv[u] = cu[63 - u] * hi;
                fv[63 - u] = cu[u] * lo;
            }
#pragma unroll
            for (int u = 0; u < 64; ++u) {
                Pb = cv[63 - u] - fv[u] * Pb;
                Qb = -fv[u] * Qb;
            }
        }
        __syncthreads();  // the forward maps are read before they are replaced
        if (tri) {
            pq[sg * NZ + m] = Pb;
            pq[(4 + sg) * NZ + m] = Qb;
        }
        __syncthreads();
        if (tri) {
            float x = 0.f;  // x_{jlo + 64}; the last segment does not use it
            for (int k = nseg - 1; k > sg; --k) x = pq[k * NZ + m] + pq[(4 + k) * NZ + m] * x;
#pragma unroll
            for (int u = 0; u < 64; ++u) {
                x = cv[63 - u] - fv[u] * x;
                col[(63 - u) * P] = x;
            }
        }
    }
    __syncthreads();

    // ---- 4. out[z][j] = E + O, out[n-1-z][j] = E - O for z < n/2, with E
    //         over the even spectral slots of the LDS block and O the odd.
    //         Wave w takes z tile w % ZT and 32 ZT columns j from (w / ZT) 32 ZT,
    //         so that it reads 1/ZT of the table (the block comes from LDS).
    auto inverse = [&](float *outp, auto acc_c, auto bf_c) {
        constexpr bool ACCUM = decltype(acc_c)::value;
        constexpr bool BFI = decltype(bf_c)::value;
        constexpr int JW = 32 * S::ZT, JT = JW / 16;
        const int zt = wave % S::ZT, jb = (wave / S::ZT) * JW;
        if (jb >= nri) return;
        f4 ev[JT], od[JT];
#pragma unroll
        for (int i = 0; i < JT; ++i) {
            ev[i] = (f4){0.f, 0.f, 0.f, 0.f};
            od[i] = (f4){0.f, 0.f, 0.f, 0.f};
        }
        // k groups [0, CT/2) are the even slots, [CT/2, CT) the odd; the
        // table fragment of the next group loads during this group's products
        auto kstep = [&](const f4 &af, int kg, f4(&acc)[JT]) {
#pragma unroll
            for (int i = 0; i < JT; ++i) {
                const f4 bf = *(const f4 *)(buf + (jb + 16 * i + t) * P + 16 * kg + 4 * q);
#pragma unroll
                for (int g = 0; g < 4; ++g) acc[i] = mma(af[g], bf[g], acc[i]);
            }
        };
        if constexpr (BFI) {
            // k groups [0, CB/2) of 32 slots are the even slots, the rest odd
            constexpr int CB = NZ / 32;
            const BH_GLOBAL bf8 *sa = gld<bf8>(a.sfa16) + (long)zt * CB * 2 * 64 + lane;
            auto kstep2 = [&](const bf8 &ah, const bf8 &al, int kb, f4(&acc)[JT]) {
#pragma unroll
                for (int i = 0; i < JT; ++i) {
                    const float *bp = buf + (jb + 16 * i + t) * P + 32 * kb + 8 * q;
                    const f4 b0 = *(const f4 *)bp, b1 = *(const f4 *)(bp + 4);
                    bf8 bh, bl;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        split_bf16(b0[e], bh, bl, e);
                        split_bf16(b1[e], bh, bl, 4 + e);
                    }
                    acc[i] = mma3(ah, al, bh, bl, acc[i]);
                }
            };
            bf8 ahc = sa[0], alc = sa[64], ahn = ahc, aln = alc;
#pragma unroll 1
            for (int kb = 0; kb < CB; ++kb) {
                if (kb + 1 < CB) {
                    ahn = sa[(kb + 1) * 2 * 64];
                    aln = sa[((kb + 1) * 2 + 1) * 64];
                }
                __builtin_amdgcn_sched_barrier(0);
                if (kb < CB / 2) kstep2(ahc, alc, kb, ev);
                else kstep2(ahc, alc, kb, od);
                ahc = ahn;
                alc = aln;
            }
        } else {
        const BH_GLOBAL f4 *sa = gld<f4>(a.sfa) + (long)zt * S::CT * 64 + lane;
        f4 afc = sa[0], afn = afc;
#pragma unroll 1
        for (int kg = 0; kg < S::CT; ++kg) {
            if (kg + 1 < S::CT) afn = sa[(kg + 1) * 64];
            __builtin_amdgcn_sched_barrier(0);
            if (kg < S::CT / 2) kstep(afc, kg, ev);
            else kstep(afc, kg, od);
            afc = afn;
        }
        }
        BH_GLOBAL float *pl = gst<float>(outp) + (long)L * numr * a.numz + numr + 1 + jb + t;
        if constexpr (ACCUM) {
            // every previous value is loaded before the first store
#pragma unroll
            for (int i = 0; i < JT; ++i) {
                if (jb + 16 * i < nri) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int z = 16 * zt + 4 * q + e;
                        const float x = ev[i][e], y = od[i][e];
                        ev[i][e] = (x + y) + pl[(long)z * numr + 16 * i];
                        od[i][e] = (x - y) + pl[(long)(NZ - 1 - z) * numr + 16 * i];
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < JT; ++i) {
                if (jb + 16 * i >= nri) break;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int z = 16 * zt + 4 * q + e;
                    pl[(long)z * numr + 16 * i] = ev[i][e];
                    pl[(long)(NZ - 1 - z) * numr + 16 * i] = od[i][e];
                }
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < JT; ++i) {
            if (jb + 16 * i >= nri) break;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int z = 16 * zt + 4 * q + e;
                const float x = ev[i][e], y = od[i][e];
                pl[(long)z * numr + 16 * i] = x + y;
                pl[(long)(NZ - 1 - z) * numr + 16 * i] = x - y;
            }
        }
    };
    // delta (split-bf16: refined later), w (FP32: never refined)
    inverse(a.dlt, std::integral_constant<bool, ACC>{}, std::integral_constant<bool, BFT>{});
    if (MR) {
        __syncthreads();  // every wave has read D
        if (fil) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int c = 0; c < CTW; ++c)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        buf[(64 * rg + 16 * i + 4 * q + e) * P + (NZ / 2) * cg + 16 * c + t] = wacc[i][c][e];
        }
        __syncthreads();
        inverse(a.w, std::false_type{}, std::false_type{});
    }
    __syncthreads();  // the block is free for the next mode
}

template <int NZ, bool MR, bool ACC, bool BFT, bool R0>
__global__ __launch_bounds__(SPEC_THREADS) void spectral_kernel(SpecArgs a, int nx, int ny, int total)
{
    extern __shared__ __align__(16) float spec_lds[];
    BH_SINGLE_LOOP(a, nx, ny, total, (spectral_tile<NZ, MR, ACC, BFT, R0>(a, spec_lds, bx)));
}

template <int NZ, bool MR, bool ACC, bool BFT, bool R0>
__global__ __launch_bounds__(SPEC_THREADS) void spectral_grouped(const SpecArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ __align__(16) float spec_lds[];
    BH_GROUP_LOOP(SpecArgs, args, gs, (spectral_tile<NZ, MR, ACC, BFT, R0>(ga, spec_lds, bx)));
}

// variant v: nzi = 32 << (v % 3); v / 3 = 0 plain, 1 with the M_r filter,
// 2 accumulating delta into dlt, 3 plain with FP32 transforms (the split
// bf16 ones wherever nzi >= 64 otherwise), 4 and 5 as 0 and 2 with a warm
// pass's fused residual from r0 (nzi >= 64)
const void *spec_fn(int variant, bool grouped)
{
#define BH_SPEC_FN(NZ, MR, ACC, BFT)                                                               \
    (grouped ? (const void *)&spectral_grouped<NZ, MR, ACC, BFT, false>                           \
             : (const void *)&spectral_kernel<NZ, MR, ACC, BFT, false>)
#define BH_SPEC_R0(NZ, ACC)                                                                        \
    (grouped ? (const void *)&spectral_grouped<NZ, false, ACC, true, true>                        \
             : (const void *)&spectral_kernel<NZ, false, ACC, true, true>)
    switch (variant) {
    case 0: return BH_SPEC_FN(32, false, false, false);
    case 1: return BH_SPEC_FN(64, false, false, true);
    case 2: return BH_SPEC_FN(128, false, false, true);
    case 3: return BH_SPEC_FN(32, true, false, false);
    case 4: return BH_SPEC_FN(64, true, false, true);
    case 5: return BH_SPEC_FN(128, true, false, true);
    case 6: return BH_SPEC_FN(32, false, true, false);
    case 7: return BH_SPEC_FN(64, false, true, true);
    case 8: return BH_SPEC_FN(128, false, true, true);
    case 9: return BH_SPEC_FN(32, false, false, false);
    case 10: return BH_SPEC_FN(64, false, false, false);
    case 11: return BH_SPEC_FN(128, false, false, false);
    case 13: return BH_SPEC_R0(64, false);
    case 14: return BH_SPEC_R0(128, false);
    case 16: return BH_SPEC_R0(64, true);
    case 17: return BH_SPEC_R0(128, true);
    default: return nullptr;
    }
#undef BH_SPEC_FN
#undef BH_SPEC_R0
}

// the block, the segment maps of the Thomas sweeps, the sub- and super-diagonal
int spec_lds_bytes(int variant)
{
    const int nz = 32 << (variant % 3);
    return (SPEC_ROWS * (nz + 4) + 8 * nz + 2 * SPEC_ROWS) * (int)sizeof(float);
}

int spec_variant(int nzi) { return nzi == 32 ? 0 : nzi == 64 ? 1 : nzi == 128 ? 2 : -1; }

}  // namespace

bool spectral_fused_ok(int nri, int nzi)
{
    const char *e = std::getenv("BH_SPEC_KERNEL");
    if (e && std::strcmp(e, "0") == 0) return false;
    return nri > 0 && nri <= SPEC_ROWS && nri % 64 == 0 && spec_variant(nzi) >= 0;
}

Op spectral_op(const SpecArgs &a, int nplanes, int nzi)
{
    const bool r0 = a.fuse_res && a.ra.r0 && !a.mr;
    return Op::make(OP_SPECTRAL,
                    spec_variant(nzi) + (a.mr ? 3 : r0 ? (a.acc ? 15 : 12) : a.acc ? 6 : a.fp32t ? 9 : 0), a,
                    nplanes, 1, 1);
}

namespace {
// bf16 bits of x, rounded to nearest even (finite x)
unsigned short bf16_bits(float x)
{
    unsigned u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
float bf16_value(unsigned short h)
{
    const unsigned u = (unsigned)h << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}
// fragment tables of nzi >= 64: [tile][32-wide k group][hi, lo][lane][8]
// bf16, packed two to a float slot (the same bytes as the FP32 tables);
// value(tile, k, lane) = S entry of lane l's element e, k = 8 (l / 16) + e
template <class F>
void split_fragments(int ntile, int nkb, F value, std::vector<float> *out)
{
    std::vector<unsigned short> h((size_t)ntile * nkb * 2 * 64 * 8);
    for (int tl = 0; tl < ntile; ++tl)
        for (int kb = 0; kb < nkb; ++kb)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const double x = value(tl, 32 * kb + 8 * (l >> 4) + e, l & 15);
                    const unsigned short hi = bf16_bits((float)x);
                    const unsigned short lo = bf16_bits((float)(x - (double)bf16_value(hi)));
                    const size_t base = (((size_t)tl * nkb + kb) * 2) * 512 + (size_t)l * 8 + e;
                    h[base] = hi;
                    h[base + 512] = lo;
                }
    out->assign(h.size() / 2, 0.f);
    std::memcpy(out->data(), h.data(), h.size() * sizeof(unsigned short));
}
}  // namespace

void spectral_fragments(int n, std::vector<float> *sfb, std::vector<float> *sfa, std::vector<float> *sfb16,
                        std::vector<float> *sfa16)
{
    const int he = n / 2, ct = n / 16, kg = he / 16, zt = he / 16;
    sfb16->clear();
    sfa16->clear();
    if (n >= 64) {
        auto sd = [&](int a, int m) { return sine_basis(a, spec_mode(m, n), n); };
        // forward B: tile = 16-column spectral tile c, k = plane a (< n/2)
        split_fragments(ct, he / 32, [&](int c, int k, int col) { return sd(k, 16 * c + col); }, sfb16);
        // inverse A: tile = 16-row z tile, k = spectral slot
        split_fragments(zt, n / 32, [&](int i, int k, int row) { return sd(16 * i + row, k); }, sfa16);
    }
    auto sf = [&](int a, int m) { return (float)sine_basis(a, spec_mode(m, n), n); };
    // B order for the forward: [c][k group][lane][4], k = plane a, n = slot m'
    sfb->assign((size_t)ct * kg * 256, 0.f);
    for (int c = 0; c < ct; ++c)
        for (int g16 = 0; g16 < kg; ++g16)
            for (int l = 0; l < 64; ++l)
                for (int g = 0; g < 4; ++g)
                    (*sfb)[(((size_t)c * kg + g16) * 64 + l) * 4 + g] = sf(16 * g16 + 4 * (l >> 4) + g, 16 * c + (l & 15));
    // A order for the inverse: [z tile][slot group][lane][4], row z, k = slot m'
    sfa->assign((size_t)zt * ct * 256, 0.f);
    for (int i = 0; i < zt; ++i)
        for (int g16 = 0; g16 < ct; ++g16)
            for (int l = 0; l < 64; ++l)
                for (int g = 0; g < 4; ++g)
                    (*sfa)[(((size_t)i * ct + g16) * 64 + l) * 4 + g] = sf(16 * i + (l & 15), 16 * g16 + 4 * (l >> 4) + g);
}

bool spectral_launch(const Op &op, hipStream_t s)
{
    if (op.kind != OP_SPECTRAL) return false;
    const void *fn = spec_fn(op.variant, false);
    if (!fn) return true;
    const int lds = spec_lds_bytes(op.variant);
    int nx = op.nx, ny = op.ny, total = (int)op.tiles();
    void *argv[] = {(void *)op.args, (void *)&nx, (void *)&ny, (void *)&total};
    const int nb = persistent_blocks(total, kernel_occupancy(fn, SPEC_THREADS, lds));
    (void)hipLaunchKernel(fn, dim3(nb), dim3(SPEC_THREADS), argv, lds, s);
    return true;
}

bool spectral_launch_grouped(int kind, int variant, const void *d_args, const GroupSched &gs, int nblocks,
                             hipStream_t s)
{
    if (kind != OP_SPECTRAL) return false;
    const void *fn = spec_fn(variant, true);
    if (!fn) return true;
    GroupSched g = gs;
    const void *ap = d_args;
    void *argv[] = {(void *)&ap, (void *)&g};
    (void)hipLaunchKernel(fn, dim3(nblocks), dim3(SPEC_THREADS), argv, spec_lds_bytes(variant), s);
    return true;
}

bool spectral_occupancy(int kind, int variant, int *blocks)
{
    if (kind != OP_SPECTRAL) return false;
    const void *fn = spec_fn(variant, true);
    *blocks = fn ? kernel_occupancy(fn, SPEC_THREADS, spec_lds_bytes(variant)) : 0;
    return true;
}

}  // namespace bh

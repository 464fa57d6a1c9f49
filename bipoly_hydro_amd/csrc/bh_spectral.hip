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
// inverse).  MFMA v_mfma_f32_16x16x4f32, operand layouts A[l%16][l/16],
// B[l/16][l%16], result element v of lane l at D[4(l/16) + v][l%16]; the
// fragment tables and M_r are stored so that lane l reads its four k values
// 4(l/16) + g of a 16-wide k group as one 16-byte load.
#include "bh_mixed.h"
#include "bh_spectral.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

namespace bh {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

constexpr int SPEC_THREADS = 512;
constexpr int SPEC_ROWS = 256;  // largest nri

__device__ __forceinline__ f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <int NZ>
struct Shape {
    static constexpr int HE = NZ / 2;     // even modes = odd modes
    static constexpr int CT = NZ / 16;    // 16-wide spectral column tiles
    static constexpr int KG = HE / 16;    // 16-wide k groups of one half
    static constexpr int ZT = HE / 16;    // 16-row z tiles of the folded inverse
    static constexpr int P = NZ + 4;      // LDS row pitch (floats)
};

template <int NZ>
__device__ __forceinline__ void spectral_tile(const SpecArgs &a, float *buf, int L)
{
    typedef Shape<NZ> S;
    constexpr int P = S::P;
    const int numr = a.numr, nri = numr - 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int t = lane & 15, q = lane >> 4;
    const int j0 = wave * 32;
    const bool active = j0 < nri;

    // ---- 1. RT = r^T S, folded: even columns from u = r_a + r_{n-1-a}, odd
    //         from v = r_a - r_{n-1-a}.  Row t of row block b is j0 + 2t + b,
    //         so one 8-byte load gives a lane both blocks' A values.
    if (active) {
        const float *r = a.r + (long)L * NZ * nri + j0 + 2 * t;
        f4 acc[2][S::CT];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < S::CT; ++c) acc[b][c] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int kg = 0; kg < S::KG; ++kg) {
            f2 u[4], v[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int pa = 16 * kg + 4 * q + g;
                const f2 x = *(const f2 *)(r + (long)pa * nri);
                const f2 y = *(const f2 *)(r + (long)(NZ - 1 - pa) * nri);
                u[g] = x + y;
                v[g] = x - y;
            }
            const f4 *sb = (const f4 *)a.sfb + kg * 64 + lane;
#pragma unroll
            for (int c = 0; c < S::CT; ++c) {
                const f4 bf = sb[c * S::KG * 64];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int b = 0; b < 2; ++b) acc[b][c] = mma(c < S::CT / 2 ? u[g][b] : v[g][b], bf[g], acc[b][c]);
            }
        }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < S::CT; ++c)
#pragma unroll
                for (int e = 0; e < 4; ++e) buf[(j0 + 2 * (4 * q + e) + b) * P + 16 * c + t] = acc[b][c][e];
    }
    __syncthreads();

    // ---- 2. W = mu . (M_r RT), C = RT - W in place; W (rows j0 + 16b + 4q + e)
    //         stays in registers for its inverse transform
    f4 wacc[2][S::CT];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < S::CT; ++c) wacc[b][c] = (f4){0.f, 0.f, 0.f, 0.f};
    if (a.mr) {
        if (active) {
            const int nkg = nri / 16;
            const f4 *mr = (const f4 *)(a.mr + (long)L * nri * nri) + lane;
            const long rstride = (long)nkg * 64;  // f4 per row tile
            f4 an[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) an[b] = mr[(2 * wave + b) * rstride];
            for (int kg = 0; kg < nkg; ++kg) {
                const f4 ac[2] = {an[0], an[1]};
                if (kg + 1 < nkg) {
#pragma unroll
                    for (int b = 0; b < 2; ++b) an[b] = mr[(2 * wave + b) * rstride + (kg + 1) * 64];
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float *row = buf + (16 * kg + 4 * q + g) * P + t;
                    float bv[S::CT];
#pragma unroll
                    for (int c = 0; c < S::CT; ++c) bv[c] = row[16 * c];
#pragma unroll
                    for (int c = 0; c < S::CT; ++c)
#pragma unroll
                        for (int b = 0; b < 2; ++b) wacc[b][c] = mma(ac[b][g], bv[c], wacc[b][c]);
                }
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int c = 0; c < S::CT; ++c) {
                const float mu = a.mu[16 * c + t];
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int idx = (j0 + 16 * b + 4 * q + e) * P + 16 * c + t;
                        const float wv = mu * wacc[b][c][e];
                        wacc[b][c][e] = wv;
                        buf[idx] = buf[idx] - wv;
                    }
            }
        }
        __syncthreads();
    }

    // ---- 3. (B_r(L) + lambda_m') D = C along j, one thread per column m'
    //         (the recurrences of the separate Thomas kernel)
    if (tid < NZ) {
        const long base = (long)L * nri * NZ + tid;
        const float *ib = a.ibet + base, *gm = a.gam + base;
        float *col = buf + tid;
        float y = 0.f;
        for (int jb = 0; jb < nri; jb += 16) {
            float iv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) iv[u] = ib[(long)(jb + u) * NZ];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int j = jb + u;
                const float aj = j > 0 ? (float)a.betar[j + 2] : 0.f;
                y = (col[j * P] - aj * y) * iv[u];
                col[j * P] = y;
            }
        }
        float x = y;
        for (int jb = nri - 2; jb >= 0; jb -= 16) {
            float gv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) gv[u] = jb - u >= 0 ? gm[(long)(jb - u + 1) * NZ] : 0.f;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int j = jb - u;
                if (j >= 0) {
                    x = col[j * P] - gv[u] * x;
                    col[j * P] = x;
                }
            }
        }
    }
    __syncthreads();

    // ---- 4. out[z][j] = E + O, out[n-1-z][j] = E - O for z < n/2, with E
    //         over the even spectral slots of the LDS block and O the odd
    auto inverse = [&](float *out) {
        if (!active) return;
        f4 ev[S::ZT][2], od[S::ZT][2];
#pragma unroll
        for (int i = 0; i < S::ZT; ++i)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                ev[i][cc] = (f4){0.f, 0.f, 0.f, 0.f};
                od[i][cc] = (f4){0.f, 0.f, 0.f, 0.f};
            }
        // k groups [0, CT/2) are the even slots, [CT/2, CT) the odd
        auto kstep = [&](int kg, f4(&acc)[S::ZT][2]) {
            f4 bf[2];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) bf[cc] = *(const f4 *)(buf + (j0 + 16 * cc + t) * P + 16 * kg + 4 * q);
            const f4 *sa = (const f4 *)a.sfa + kg * 64 + lane;
#pragma unroll
            for (int i = 0; i < S::ZT; ++i) {
                const f4 af = sa[i * S::CT * 64];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc) acc[i][cc] = mma(af[g], bf[cc][g], acc[i][cc]);
            }
        };
#pragma unroll 1
        for (int kg = 0; kg < S::CT / 2; ++kg) kstep(kg, ev);
#pragma unroll 1
        for (int kg = S::CT / 2; kg < S::CT; ++kg) kstep(kg, od);
        float *pl = out + (long)L * numr * a.numz + numr + 1 + j0 + t;
#pragma unroll
        for (int i = 0; i < S::ZT; ++i)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int z = 16 * i + 4 * q + e;
                    const float x = ev[i][cc][e], y = od[i][cc][e];
                    pl[(long)z * numr + 16 * cc] = x + y;
                    pl[(long)(NZ - 1 - z) * numr + 16 * cc] = x - y;
                }
    };
    inverse(a.dlt);
    if (a.mr) {
        __syncthreads();  // every wave has read D
        if (active) {
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < S::CT; ++c)
#pragma unroll
                    for (int e = 0; e < 4; ++e) buf[(j0 + 16 * b + 4 * q + e) * P + 16 * c + t] = wacc[b][c][e];
        }
        __syncthreads();
        inverse(a.w);
    }
    __syncthreads();  // the block is free for the next mode
}

template <int NZ>
__global__ __launch_bounds__(SPEC_THREADS) void spectral_kernel(SpecArgs a, int nx, int ny, int total)
{
    extern __shared__ __align__(16) float spec_lds[];
    BH_SINGLE_LOOP(a, nx, ny, total, (spectral_tile<NZ>(a, spec_lds, bx)));
}

template <int NZ>
__global__ __launch_bounds__(SPEC_THREADS) void spectral_grouped(const SpecArgs *__restrict__ args, GroupSched gs)
{
    extern __shared__ __align__(16) float spec_lds[];
    BH_GROUP_LOOP(SpecArgs, args, gs, (spectral_tile<NZ>(ga, spec_lds, bx)));
}

// variant v: nzi = 32 << v
const void *spec_fn(int variant, bool grouped)
{
    switch (variant) {
    case 0: return grouped ? (const void *)&spectral_grouped<32> : (const void *)&spectral_kernel<32>;
    case 1: return grouped ? (const void *)&spectral_grouped<64> : (const void *)&spectral_kernel<64>;
    case 2: return grouped ? (const void *)&spectral_grouped<128> : (const void *)&spectral_kernel<128>;
    default: return nullptr;
    }
}

int spec_lds_bytes(int variant) { return SPEC_ROWS * ((32 << variant) + 4) * (int)sizeof(float); }

int spec_variant(int nzi) { return nzi == 32 ? 0 : nzi == 64 ? 1 : nzi == 128 ? 2 : -1; }

}  // namespace

bool spectral_fused_ok(int nri, int nzi)
{
    const char *e = std::getenv("BH_SPEC_KERNEL");
    if (e && std::strcmp(e, "0") == 0) return false;
    return nri > 0 && nri <= SPEC_ROWS && nri % 32 == 0 && spec_variant(nzi) >= 0;
}

Op spectral_op(const SpecArgs &a, int nplanes, int nzi)
{
    return Op::make(OP_SPECTRAL, spec_variant(nzi), a, nplanes, 1, 1);
}

void spectral_fragments(int n, std::vector<float> *sfb, std::vector<float> *sfa)
{
    const int he = n / 2, ct = n / 16, kg = he / 16, zt = he / 16;
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

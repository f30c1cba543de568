// rad_gather.hip - dense form-factor matrix and the Jacobi gather on MFMA.
//
// The shooting update of the reference (Main.cpp:1251-1279) distributes one
// shooter's unshot radiosity along its hemicube row F_i.  Collecting every
// row gives the dense matrix F (F_ij = form factor from patch i to patch j),
// and the same fixed point is reached by gathering:
//     B = E + diag(rho) F B                      (per colour channel)
// The kernels here keep F in HBM as BF16 (round to nearest even) plus,
// optionally, the BF16 remainder F - bf16(F), and apply it to up to 192
// right-hand-side columns with v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
//
// Layouts (npad = N rounded up to 512, padding is zero):
//   F      npad x npad u16 (bf16 bits), row i = hemicube row of patch i
//   X      rp x npad u16: column r of the right-hand side, contiguous over j
//   vectors (E, B, r) rp x npad f32, same column-major order
//   YP     S x rp x npad f32: split-K partial products F X
#include "rad_internal.h"

namespace rad {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Element (r, c) of a dense npad x npad half (fhi or flo).  The halves are
// stored in 4 KB tiles of 32 rows x 64 columns, tile (r / 32, c / 64) at
// element ((r / 32) * (npad / 64) + c / 64) * 2048, so a GEMM wavefront's
// 32 rows of one 64-column chunk are one contiguous 4 KB block and its row
// tile one contiguous stream.  Inside a tile the 16-byte units of 8 columns
// are in the order of the MFMA A fragments: unit (f * 2 + ks) * 64 + lane
// for rows 16 f .. 16 f + 15 and columns 32 ks .. 32 ks + 31 of the tile,
// lane = (column / 8 % 4) * 16 + row % 16.  Eight consecutive columns from a
// multiple of 8 are one unit.
__device__ __forceinline__ size_t dense_off(uint32_t r, uint32_t c, uint32_t npad)
{
    const uint32_t rr = r & 31u, cc = c & 63u;
    const uint32_t unit = ((rr >> 4) * 2u + (cc >> 5)) * 64u + ((cc >> 3) & 3u) * 16u + (rr & 15u);
    return ((size_t)(r >> 5) * (npad >> 6) + (c >> 6)) * 2048u + unit * 8u + (cc & 7u);
}

// ---------------------------------------------------------------- F build

// Slots [0,k) of a pass render the hemicubes of patches row0 .. row0+k-1.
__global__ void k_iota_shooters(uint32_t row0, uint32_t k, uint32_t *shooters, uint32_t *nshoot)
{
    uint32_t s = threadIdx.x + blockIdx.x * blockDim.x;
    if (s < k)
        shooters[s] = row0 + s;
    if (s == 0)
        *nshoot = k;
}

// The next pass of a captured dense build: ctl[kCtlDenseBase] advances by K
// (one workgroup of at least K threads) and its rows become the shooters.
__global__ void k_iota_next(uint32_t K, uint32_t npatch, uint32_t *__restrict__ ctl,
                            uint32_t *__restrict__ shooters, uint32_t *__restrict__ nshoot)
{
    __shared__ uint32_t base;
    if (threadIdx.x == 0) {
        base = ctl[kCtlDenseBase] + K;
        ctl[kCtlDenseBase] = base;
    }
    __syncthreads();
    const uint32_t k = base < npatch ? min(K, npatch - base) : 0u;
    if (threadIdx.x < k)
        shooters[threadIdx.x] = base + threadIdx.x;
    if (threadIdx.x == 0)
        *nshoot = k;
}

// Rows row0 .. row0+k-1 of F from the F rows of one pass (k x npatch f32).
// One lane writes 8 consecutive columns (16 B of hi, 16 B of lo).
__global__ void k_dense_store(const float *__restrict__ ff, uint32_t k, uint32_t row0,
                              uint32_t npatch, uint32_t npad, uint16_t *__restrict__ fhi,
                              uint16_t *__restrict__ flo)
{
    uint32_t s = blockIdx.y;
    if (s >= k)
        return;
    const float *src = ff + (size_t)s * npatch;
    const uint32_t row = row0 + s;
    for (uint32_t j0 = (blockIdx.x * blockDim.x + threadIdx.x) * 8; j0 < npad;
         j0 += gridDim.x * blockDim.x * 8) {
        uint16_t h[8], l[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            uint32_t j = j0 + q;
            float f = j < npatch ? src[j] : 0.0f;
            h[q] = bf16_rne(f);
            l[q] = bf16_rne(f - bf16_val(h[q]));
        }
        uint4 vh, vl;
        vh.x = h[0] | ((uint32_t)h[1] << 16);
        vh.y = h[2] | ((uint32_t)h[3] << 16);
        vh.z = h[4] | ((uint32_t)h[5] << 16);
        vh.w = h[6] | ((uint32_t)h[7] << 16);
        vl.x = l[0] | ((uint32_t)l[1] << 16);
        vl.y = l[2] | ((uint32_t)l[3] << 16);
        vl.z = l[4] | ((uint32_t)l[5] << 16);
        vl.w = l[6] | ((uint32_t)l[7] << 16);
        const size_t o = dense_off(row, j0, npad);
        *(uint4 *)(fhi + o) = vh;
        if (flo)
            *(uint4 *)(flo + o) = vl;
    }
}

// Shooting-consistent layout: row j of G holds G_ji = F_ij A_i / A_j, the
// factor PHYSICAL shooting applies when shooter i sends to receiver j
// (k_shoot), so B = E + diag(rho) G B is the fixed point of the shooting
// iteration.  Pass rows row0 .. row0+k-1 of F become columns of G.  A block
// takes 64 receivers; for each chunk of 64 shooters it reads the F entries
// along the rows (coalesced), scales and splits them, transposes them
// through LDS and writes 128-byte runs of each G row (8 lanes x 16 B of hi
// and of lo).  Only the k columns of this pass are written.  Each thread
// issues its 16 loads of a chunk at once, and sets every nonzero F entry it
// read back to zero, so the pass's F rows are left zero for the next one.
// With dctl != null (a captured build), row0 is dctl[kCtlDenseBase] and k
// is *dk.
__global__ __launch_bounds__(256) void k_dense_store_t(float *__restrict__ ff, uint32_t k,
                                                       uint32_t row0, uint32_t npatch,
                                                       uint32_t npad,
                                                       const float *__restrict__ area,
                                                       uint16_t *__restrict__ fhi,
                                                       uint16_t *__restrict__ flo,
                                                       const uint32_t *__restrict__ dctl,
                                                       const uint32_t *__restrict__ dk)
{
    if (dctl) {
        row0 = dctl[kCtlDenseBase];
        k = *dk;
    }
    __shared__ uint32_t tile[64][65]; // hi | lo << 16 per (receiver, shooter)
    const uint32_t t = threadIdx.x, jl = t & 63;
    const uint32_t j0 = blockIdx.x * 64, j = j0 + jl;
    const bool jok = j < npatch;
    const float inv_aj = jok ? 1.0f / area[j] : 0.0f;
    for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        const uint32_t nc = min(64u, k - c0);
        float fv[16], as[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t sl = (t >> 6) + 4u * (uint32_t)q;
            const bool ok = jok && sl < nc;
            fv[q] = ok ? ff[(size_t)(c0 + sl) * npatch + j] : 0.0f;
            as[q] = sl < nc ? area[row0 + c0 + sl] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t sl = (t >> 6) + 4u * (uint32_t)q;
            if (fv[q] != 0.0f)
                ff[(size_t)(c0 + sl) * npatch + j] = 0.0f;
            const float f = fv[q] * as[q] * inv_aj;
            const uint16_t h = bf16_rne(f), l = bf16_rne(f - bf16_val(h));
            tile[jl][sl] = (uint32_t)h | ((uint32_t)l << 16);
        }
        __syncthreads();
        const uint32_t s8 = 8 * (t & 7);
        for (uint32_t r = t >> 3; r < 64; r += 32) {
            const uint32_t jj = j0 + r;
            if (jj >= npatch || s8 >= nc)
                continue;
            const uint32_t col = row0 + c0 + s8;
            uint32_t w[8];
#pragma unroll
            for (int e = 0; e < 8; ++e)
                w[e] = tile[r][s8 + e];
            if (((row0 + c0) & 7u) == 0 && s8 + 8 <= nc) {
                uint4 vh, vl;
                vh.x = (w[0] & 0xFFFFu) | (w[1] << 16);
                vh.y = (w[2] & 0xFFFFu) | (w[3] << 16);
                vh.z = (w[4] & 0xFFFFu) | (w[5] << 16);
                vh.w = (w[6] & 0xFFFFu) | (w[7] << 16);
                vl.x = (w[0] >> 16) | (w[1] & 0xFFFF0000u);
                vl.y = (w[2] >> 16) | (w[3] & 0xFFFF0000u);
                vl.z = (w[4] >> 16) | (w[5] & 0xFFFF0000u);
                vl.w = (w[6] >> 16) | (w[7] & 0xFFFF0000u);
                const size_t o = dense_off(jj, col, npad);
                *(uint4 *)(fhi + o) = vh;
                *(uint4 *)(flo + o) = vl;
            } else {
                for (uint32_t e = 0; e < 8 && s8 + e < nc; ++e) {
                    const size_t o = dense_off(jj, col + e, npad);
                    fhi[o] = (uint16_t)(w[e] & 0xFFFFu);
                    flo[o] = (uint16_t)(w[e] >> 16);
                }
            }
        }
        __syncthreads();
    }
}

// Row-major f32 matrix (rows x npatch, host order) into rows of F.
__global__ void k_dense_from_f32(const float *__restrict__ src, uint32_t rows, uint32_t row0,
                                 uint32_t npatch, uint32_t npad, uint16_t *__restrict__ fhi,
                                 uint16_t *__restrict__ flo)
{
    size_t n = (size_t)rows * npatch;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        size_t i = e / npatch, j = e - i * npatch;
        float f = src[e];
        uint16_t h = bf16_rne(f);
        const size_t o = dense_off((uint32_t)(row0 + i), (uint32_t)j, npad);
        fhi[o] = h;
        if (flo)
            flo[o] = bf16_rne(f - bf16_val(h));
    }
}

__global__ void k_dense_to_f32(const uint16_t *__restrict__ fhi, const uint16_t *__restrict__ flo,
                               uint32_t rows, uint32_t row0, uint32_t npatch, uint32_t npad,
                               float *__restrict__ dst)
{
    size_t n = (size_t)rows * npatch;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        size_t i = e / npatch, j = e - i * npatch;
        const size_t o = dense_off((uint32_t)(row0 + i), (uint32_t)j, npad);
        float f = bf16_val(fhi[o]);
        if (flo)
            f += bf16_val(flo[o]);
        dst[e] = f;
    }
}

// ------------------------------------------------------------------ GEMM

// One 16-byte A fragment load: non-temporal when the product has several
// column tiles (the matrix is read once, X should stay in L2).
template <int RT> __device__ __forceinline__ uint4 gemm_aload(const uint16_t *p)
{
    if constexpr (RT > 1) {
        u32x4 v = __builtin_nontemporal_load((const u32x4 *)p);
        return uint4{v.x, v.y, v.z, v.w};
    } else {
        return *(const uint4 *)p;
    }
}

// YP[s] = F[:, k slice s] X[k slice s]  for 32 rows per wave.
// Wave w owns rows i0 .. i0+31 (two 16-row A fragments) and all RT column
// tiles of 16.  Per 64-wide k chunk the X slice (RT*16 rows x 64 bf16) is
// staged in LDS and shared by the waves; the A fragments go straight
// from HBM to registers (each F element is used by one wave only), one
// contiguous 4 KB tile of the matrix per wave and chunk (dense_off).  PF
// chunks of A fragments are in flight ahead of the one being multiplied
// (PF = 2: two register sets, the loop unrolled by two chunks), so a load
// has PF chunks of MFMA work to arrive in.
// SPLIT accumulates Fhi Xhi + Fhi Xlo + Flo Xhi (about 16 significant bits
// of F and X) for the refinement residual.
template <int RT, bool SPLIT, int WAVES, int PF>
__global__ __launch_bounds__(WAVES * 64) void k_gemm(const uint16_t *__restrict__ fhi,
                                              const uint16_t *__restrict__ flo,
                                              const uint16_t *__restrict__ xhi,
                                              const uint16_t *__restrict__ xlo, uint32_t npad,
                                              GemmSplit ks, float *__restrict__ yp)
{
    static_assert(PF == 1 || PF == 2, "one or two chunks of A in flight");
    constexpr int NB = SPLIT ? 2 : 1;
    constexpr int LROW = 72;               // bf16 per LDS row: 64 + 8 pad (144 B)
    constexpr int UNITS = RT * 16 * 8;     // 16-byte units of one X chunk
    constexpr int NT = WAVES * 64;
    constexpr int UPT = (UNITS + NT - 1) / NT;
    // every thread stages UPT units when NT divides UNITS: no guard, so the
    // staging registers are not demoted to scratch memory
    constexpr bool FULL = UNITS % NT == 0;
    // two X buffers when they fit (one product half): chunk c + 1 is staged
    // while other waves still read chunk c, one barrier per chunk
    constexpr int DB = NB == 1 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) uint16_t lds[DB][NB][RT * 16][LROW];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kg = lane >> 4;
    // 1-D grid, k slice fastest: consecutive workgroups go to different XCDs
    // (round robin), so with S dividing 8 every XCD works on one k slice and
    // keeps that slice of X in its own L2.  The row tiles from ks.rsplit on
    // come last, in ks.S2 slices.
    const uint32_t head = ks.rsplit / (WAVES * 32) * ks.S;
    const bool tail = blockIdx.x >= head;
    const uint32_t ns = tail ? ks.S2 : ks.S, b = tail ? blockIdx.x - head : blockIdx.x;
    const uint32_t slice = b % ns;
    const uint32_t rtile = (tail ? ks.rsplit / (WAVES * 32) : 0u) + b / ns;
    const uint32_t i0 = rtile * (WAVES * 32) + wave * 32;
    // the last row tile may be partial (npad is a multiple of 512, a tile
    // of 12 wavefronts holds 384 rows): its waves beyond npad only stage
    // their share of X
    const bool live = i0 < npad;
    const uint32_t nch = npad >> 6;
    const uint32_t kbeg = slice * nch / ns * 64u, kend = (slice + 1u) * nch / ns * 64u;

    f32x4 acc[2][RT];
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
        for (int t = 0; t < RT; ++t)
            acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // the wave's row tile: chunk kc is the 4 KB block kc / 64 of it, the
    // lane's 16-byte unit of fragment (f, ks) at (f * 2 + ks) * 512 + 8 lane
    const size_t abase = (size_t)(i0 >> 5) * (npad >> 6) * 2048u + (size_t)lane * 8u;

    static_assert(UPT <= 3, "three staging units per thread at most");
    // staging registers of the X chunk units tid, tid + NT, tid + 2 NT (hi,
    // lo): named variables, which stay in VGPRs where an array was kept in
    // scratch memory
    uint4 xh0, xh1, xh2, xl0, xl1, xl2;
    uint4 ar[NB][2][2], br[NB][2][2], ac[NB][2][2];

    // X chunk kc into registers (staged to LDS at the top of the chunk's
    // iteration) and the A fragments of chunk kc.  The matrix is streamed
    // once per product: with several column tiles non-temporal loads keep it
    // from evicting X out of L2.
#define RAD_GEMM_XUNIT(u_, h_, l_, kc_)                                                \
    if (FULL || tid + (u_) * NT < UNITS) {                                             \
        const int idx = tid + (u_) * NT;                                               \
        const size_t o = (size_t)(idx >> 3) * npad + (kc_) + (idx & 7) * 8;            \
        h_ = *(const uint4 *)(xhi + o);                                                \
        if (SPLIT)                                                                     \
            l_ = *(const uint4 *)(xlo + o);                                            \
    }
#define RAD_GEMM_LOAD_X(kc_)                                                           \
    {                                                                                  \
        RAD_GEMM_XUNIT(0, xh0, xl0, kc_)                                               \
        if constexpr (UPT > 1) {                                                       \
            RAD_GEMM_XUNIT(1, xh1, xl1, kc_)                                           \
        }                                                                              \
        if constexpr (UPT > 2) {                                                       \
            RAD_GEMM_XUNIT(2, xh2, xl2, kc_)                                           \
        }                                                                              \
    }
#define RAD_GEMM_LOAD_A(dst_, kc_)                                                     \
    if (live)                                                                          \
    _Pragma("unroll") for (int ks = 0; ks < 2; ++ks)                                   \
    {                                                                                  \
        const size_t a0_ = abase + (size_t)((kc_) >> 6) * 2048u + ks * 512;            \
        dst_[0][0][ks] = gemm_aload<RT>(fhi + a0_);                                    \
        dst_[0][1][ks] = gemm_aload<RT>(fhi + a0_ + 1024);                             \
        if (SPLIT) {                                                                   \
            dst_[NB - 1][0][ks] = gemm_aload<RT>(flo + a0_);                           \
            dst_[NB - 1][1][ks] = gemm_aload<RT>(flo + a0_ + 1024);                    \
        }                                                                              \
    }
    // X of the chunk to LDS buffer cb_ (after the barrier that frees it)
#define RAD_GEMM_STAGE_X(cb_)                                                          \
    {                                                                                  \
        if (DB == 1)                                                                   \
            __syncthreads();                                                           \
        _Pragma("unroll") for (int u = 0; u < UPT; ++u)                                \
        {                                                                              \
            int idx = tid + u * NT;                                                    \
            if (FULL || idx < UNITS) {                                                 \
                const uint4 h = u == 0 ? xh0 : u == 1 ? xh1 : xh2;                     \
                *(uint4 *)&lds[cb_][0][idx >> 3][(idx & 7) * 8] = h;                   \
                if (SPLIT) {                                                           \
                    const uint4 l = u == 0 ? xl0 : u == 1 ? xl1 : xl2;                 \
                    *(uint4 *)&lds[cb_][NB - 1][idx >> 3][(idx & 7) * 8] = l;          \
                }                                                                      \
            }                                                                          \
        }                                                                              \
        __syncthreads();                                                               \
    }
#define RAD_GEMM_TAKE(src_)                                                            \
    _Pragma("unroll") for (int b = 0; b < NB; ++b)                                     \
        _Pragma("unroll") for (int f = 0; f < 2; ++f)                                  \
            _Pragma("unroll") for (int ks = 0; ks < 2; ++ks)                           \
                ac[b][f][ks] = src_[b][f][ks];
#define RAD_GEMM_MFMA(cb_)                                                             \
    if (live)                                                                          \
    _Pragma("unroll") for (int ks = 0; ks < 2; ++ks)                                   \
    {                                                                                  \
        _Pragma("unroll") for (int t = 0; t < RT; ++t)                                 \
        {                                                                              \
            bf16x8 bh = *(const bf16x8 *)&lds[cb_][0][t * 16 + m][ks * 32 + 8 * kg];   \
            _Pragma("unroll") for (int f = 0; f < 2; ++f)                              \
            {                                                                          \
                bf16x8 ah = __builtin_bit_cast(bf16x8, ac[0][f][ks]);                  \
                acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[f][t], 0, 0, 0); \
                if (SPLIT) {                                                           \
                    bf16x8 bl = *(const bf16x8 *)&lds[cb_][NB - 1][t * 16 + m][ks * 32 + 8 * kg]; \
                    bf16x8 al = __builtin_bit_cast(bf16x8, ac[NB - 1][f][ks]);         \
                    acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[f][t], 0, 0, 0); \
                    acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[f][t], 0, 0, 0); \
                }                                                                      \
            }                                                                          \
        }                                                                              \
    }

    RAD_GEMM_LOAD_X(kbeg)
    RAD_GEMM_LOAD_A(ar, kbeg)
    if constexpr (PF == 1) {
        for (uint32_t kc = kbeg, ci = 0; kc < kend; kc += 64, ++ci) {
            const int cb = DB == 2 ? (int)(ci & 1u) : 0;
            // the next chunk's A loads go out before the barrier, so they
            // are in flight while the wave waits for the others
            RAD_GEMM_TAKE(ar)
            if (kc + 64 < kend)
                RAD_GEMM_LOAD_A(ar, kc + 64)
            RAD_GEMM_STAGE_X(cb)
            if (kc + 64 < kend)
                RAD_GEMM_LOAD_X(kc + 64)
            RAD_GEMM_MFMA(cb)
        }
    } else {
        if (kbeg + 64 < kend)
            RAD_GEMM_LOAD_A(br, kbeg + 64)
        // chunks in pairs: the first of a pair in ar and LDS buffer 0, the
        // second in br and buffer DB - 1; each chunk's X is loaded one chunk
        // ahead, its A two chunks ahead
        // (scheduling barriers keep the two halves apart: merged, their LDS
        // reads were hoisted together and the registers spilled)
        for (uint32_t kc = kbeg; kc < kend; kc += 128) {
            RAD_GEMM_STAGE_X(0)
            RAD_GEMM_TAKE(ar)
            if (kc + 64 < kend)
                RAD_GEMM_LOAD_X(kc + 64)
            if (kc + 128 < kend)
                RAD_GEMM_LOAD_A(ar, kc + 128)
            RAD_GEMM_MFMA(0)
            __builtin_amdgcn_sched_barrier(0);
            if (kc + 64 >= kend)
                break;
            RAD_GEMM_STAGE_X(DB - 1)
            RAD_GEMM_TAKE(br)
            if (kc + 128 < kend)
                RAD_GEMM_LOAD_X(kc + 128)
            if (kc + 192 < kend)
                RAD_GEMM_LOAD_A(br, kc + 192)
            RAD_GEMM_MFMA(DB - 1)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // lane holds rows 4*kg .. 4*kg+3 of column m of each 16x16 tile
    float *out = yp + (size_t)slice * (RT * 16) * npad;
    if (!live)
        return;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
        for (int t = 0; t < RT; ++t)
            *(f32x4 *)(out + (size_t)(t * 16 + m) * npad + i0 + f * 16 + 4 * kg) = acc[f][t];
}
#undef RAD_GEMM_LOAD_X
#undef RAD_GEMM_XUNIT
#undef RAD_GEMM_LOAD_A
#undef RAD_GEMM_STAGE_X
#undef RAD_GEMM_TAKE
#undef RAD_GEMM_MFMA

// ------------------------------------------------------- vector updates

// Block partial sums in FP64 (one double per block).
__device__ __forceinline__ void block_sum_store(double v, double *parts)
{
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        parts[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Update kinds of k_gather_update (one lane per (column, patch) element):
//  kInit    r = E, B = 0, X = bf16(r); partial sums of A|E|
//  kInner   d = r + rho y, X = bf16(d)
//  kLast    d = r + rho y, B += d, (Xhi, Xlo) = split(B)
//  kResid   r = E + rho y - B (FP64), X = bf16(r); partial sums of A|r|
//  kOut     out = B + r
enum { kInit = 0, kInner = 1, kLast = 2, kResid = 3 };

__global__ __launch_bounds__(256) void k_gather_update(
    int kind, const float *__restrict__ yp, GemmSplit ks, uint32_t rp, uint32_t R, uint32_t npad,
    uint32_t n, const float *__restrict__ E, const float *__restrict__ refl,
    const float *__restrict__ area, float *__restrict__ rv, float *__restrict__ B,
    uint16_t *__restrict__ xhi, uint16_t *__restrict__ xlo, double *__restrict__ parts)
{
    double acc = 0.0;
    size_t total = (size_t)R * npad;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t plane = (size_t)rp * npad;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += stride) {
        uint32_t r = (uint32_t)(e / npad), i = (uint32_t)(e - (size_t)r * npad);
        if (i >= n)
            continue;
        if (kind == kInit) {
            float v = E[e];
            rv[e] = v;
            B[e] = 0.0f;
            xhi[e] = bf16_rne(v);
            acc += (double)area[i] * fabs((double)v);
            continue;
        }
        float y = 0.0f;
        const uint32_t S = i < ks.rsplit ? ks.S : ks.S2;
        for (uint32_t s = 0; s < S; ++s)
            y += yp[s * plane + e];
        float rho = refl[(size_t)i * 3 + r % 3];
        if (kind == kResid) {
            double res = (double)E[e] + (double)rho * (double)y - (double)B[e];
            float rf = (float)res;
            rv[e] = rf;
            xhi[e] = bf16_rne(rf);
            acc += (double)area[i] * fabs(res);
        } else {
            float d = rv[e] + rho * y;
            if (kind == kInner) {
                xhi[e] = bf16_rne(d);
            } else {
                float b = B[e] + d;
                B[e] = b;
                uint16_t h = bf16_rne(b);
                xhi[e] = h;
                xlo[e] = bf16_rne(b - bf16_val(h));
            }
        }
    }
    if (kind == kInit || kind == kResid)
        block_sum_store(acc, parts);
}

__global__ void k_gather_out(uint32_t R, uint32_t npad, uint32_t n, const float *__restrict__ B,
                             const float *__restrict__ rv, float *__restrict__ out)
{
    size_t total = (size_t)R * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        size_t r = e / n, i = e - r * n;
        size_t o = r * npad + i;
        out[e] = (float)((double)B[o] + (double)rv[o]);
    }
}

// Columns in from host order (R x n) into the padded device layout.
__global__ void k_pad_columns(uint32_t R, uint32_t npad, uint32_t n, const float *__restrict__ src,
                              float *__restrict__ dst, uint16_t *__restrict__ xhi,
                              uint16_t *__restrict__ xlo)
{
    size_t total = (size_t)R * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        size_t r = e / n, i = e - r * n;
        size_t o = r * npad + i;
        float v = src[e];
        if (dst)
            dst[o] = v;
        if (xhi) {
            uint16_t h = bf16_rne(v);
            xhi[o] = h;
            if (xlo)
                xlo[o] = bf16_rne(v - bf16_val(h));
        }
    }
}

// Sum of the S partial products into host column order (R x n).
__global__ void k_sum_partials(uint32_t R, uint32_t npad, uint32_t n, GemmSplit ks, uint32_t rp,
                               const float *__restrict__ yp, float *__restrict__ out)
{
    size_t total = (size_t)R * n;
    size_t plane = (size_t)rp * npad;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        size_t r = e / n, i = e - r * n;
        float y = 0.0f;
        const uint32_t S = i < ks.rsplit ? ks.S : ks.S2;
        for (uint32_t s = 0; s < S; ++s)
            y += yp[s * plane + r * npad + i];
        out[e] = y;
    }
}

// ---------------------------------------------------------- FP64 checks
// The FP64 residual of a gather solution against freshly rendered FP32
// hemicube rows (shooting layout M_ji = F_ij A_i / A_j), streamed one pass
// of k rows at a time:
//  k_check_accum   acc[r][j] += sum_s F[s][j] * A_i B[r][i]  (i = row0 + s),
//                  one lane per receiver j, FP64
//  k_check_rowsum  q = max over rows and channels of sum_j rho_jc F_ij
//                  (one workgroup per row; positive doubles ordered as u64)
//  k_check_final   r = E + rho acc / A - B, partial sums of A|r| and A|B|
constexpr uint32_t kCheckMaxR = 16;

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_check_accum(const float *__restrict__ ff, uint32_t k,
                                                     uint32_t row0, uint32_t rowlen, uint32_t n,
                                                     uint32_t R, const float *__restrict__ B,
                                                     const float *__restrict__ area,
                                                     double *__restrict__ acc)
{
    __shared__ double w[kCheckMaxR][RAD_MAX_K];
    for (uint32_t e = threadIdx.x; e < R * k; e += blockDim.x) {
        const uint32_t r = e / k, s = e - r * k, i = row0 + s;
        w[r][s] = (double)area[i] * (double)B[(size_t)r * n + i];
    }
    __syncthreads();
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    double a[kCheckMaxR];
    for (uint32_t r = 0; r < R; ++r)
        a[r] = 0.0;
    for (uint32_t s = 0; s < k; ++s) {
        const float f = ff[(size_t)s * rowlen + j];
        if (f == 0.0f)
            continue;
        for (uint32_t r = 0; r < R; ++r)
            a[r] += (double)f * w[r][s];
    }
    for (uint32_t r = 0; r < R; ++r)
        acc[(size_t)r * n + j] += a[r];
}

__global__ __launch_bounds__(1024) void k_check_rowsum(const float *__restrict__ ff, uint32_t rowlen,
                                                       uint32_t n, const float *__restrict__ refl,
                                                       unsigned long long *__restrict__ qbits)
{
    __shared__ double part[3][16];
    const uint32_t s = blockIdx.x;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const double f = (double)ff[(size_t)s * rowlen + j];
        c0 += f * (double)refl[(size_t)j * 3 + 0];
        c1 += f * (double)refl[(size_t)j * 3 + 1];
        c2 += f * (double)refl[(size_t)j * 3 + 2];
    }
    c0 = wave_sum_d(c0);
    c1 = wave_sum_d(c1);
    c2 = wave_sum_d(c2);
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wv] = c0;
        part[1][wv] = c1;
        part[2][wv] = c2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
            for (uint32_t q = 0; q < (blockDim.x >> 6); ++q)
                t += part[c][q];
            m = t > m ? t : m;
        }
        atomicMax(qbits, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ __launch_bounds__(256) void k_check_final(uint32_t R, uint32_t n,
                                                     const double *__restrict__ acc,
                                                     const float *__restrict__ E,
                                                     const float *__restrict__ B,
                                                     const float *__restrict__ refl,
                                                     const float *__restrict__ area,
                                                     double *__restrict__ parts)
{
    double num = 0.0, den = 0.0;
    const size_t total = (size_t)R * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(e / n), j = (uint32_t)(e - (size_t)r * n);
        const double a = (double)area[j];
        const double res = (double)E[e] + (double)refl[(size_t)j * 3 + r % 3] * acc[e] / a -
                           (double)B[e];
        num += a * fabs(res);
        den += a * fabs((double)B[e]);
    }
    block_sum_store(num, parts);
    __syncthreads(); // block_sum_store's LDS is reused
    block_sum_store(den, parts + gridDim.x);
}

void launch_check_accum(const float *ff, uint32_t k, uint32_t row0, uint32_t rowlen, uint32_t n,
                        uint32_t R, const float *B, const float *area, double *acc,
                        hipStream_t st)
{
    k_check_accum<<<(n + 255) / 256, 256, 0, st>>>(ff, k, row0, rowlen, n, R, B, area, acc);
}

void launch_check_rowsum(const float *ff, uint32_t k, uint32_t rowlen, uint32_t n,
                         const float *refl, unsigned long long *qbits, hipStream_t st)
{
    k_check_rowsum<<<k, 1024, 0, st>>>(ff, rowlen, n, refl, qbits);
}

void launch_check_final(uint32_t R, uint32_t n, const double *acc, const float *E, const float *B,
                        const float *refl, const float *area, double *parts, uint32_t nblocks,
                        hipStream_t st)
{
    k_check_final<<<nblocks, 256, 0, st>>>(R, n, acc, E, B, refl, area, parts);
}

// --------------------------------------------------------------- launchers

void launch_iota_shooters(uint32_t row0, uint32_t k, uint32_t *shooters, uint32_t *nshoot,
                          hipStream_t st)
{
    k_iota_shooters<<<(k + 255) / 256, 256, 0, st>>>(row0, k, shooters, nshoot);
}

void launch_dense_store(const float *ff, uint32_t k, uint32_t row0, uint32_t npatch, uint32_t npad,
                        uint16_t *fhi, uint16_t *flo, hipStream_t st)
{
    uint32_t bx = (npad / 8 + 255) / 256;
    if (bx > 64)
        bx = 64;
    k_dense_store<<<dim3(bx, k), 256, 0, st>>>(ff, k, row0, npatch, npad, fhi, flo);
}

void launch_dense_store_t(float *ff, uint32_t k, uint32_t row0, uint32_t npatch,
                          uint32_t npad, const float *area, uint16_t *fhi, uint16_t *flo,
                          hipStream_t st)
{
    k_dense_store_t<<<(npatch + 63) / 64, 256, 0, st>>>(ff, k, row0, npatch, npad, area, fhi,
                                                        flo, nullptr, nullptr);
}

void launch_iota_next(uint32_t K, uint32_t npatch, uint32_t *ctl, uint32_t *shooters,
                      uint32_t *nshoot, hipStream_t st)
{
    k_iota_next<<<1, (K + 63) / 64 * 64, 0, st>>>(K, npatch, ctl, shooters, nshoot);
}

void launch_dense_store_t_dev(float *ff, uint32_t K, const uint32_t *ctl, const uint32_t *nshoot,
                              uint32_t npatch, uint32_t npad, const float *area, uint16_t *fhi,
                              uint16_t *flo, hipStream_t st)
{
    k_dense_store_t<<<(npatch + 63) / 64, 256, 0, st>>>(ff, K, 0u, npatch, npad, area, fhi, flo,
                                                        ctl, nshoot);
}

void launch_dense_from_f32(const float *src, uint32_t rows, uint32_t row0, uint32_t npatch,
                           uint32_t npad, uint16_t *fhi, uint16_t *flo, hipStream_t st)
{
    k_dense_from_f32<<<2048, 256, 0, st>>>(src, rows, row0, npatch, npad, fhi, flo);
}

void launch_dense_to_f32(const uint16_t *fhi, const uint16_t *flo, uint32_t rows, uint32_t row0,
                         uint32_t npatch, uint32_t npad, float *dst, hipStream_t st)
{
    k_dense_to_f32<<<2048, 256, 0, st>>>(fhi, flo, rows, row0, npatch, npad, dst);
}

uint32_t gemm_tiles(uint32_t R)
{
    uint32_t t = (R + 15) / 16;
    if (t <= 1)
        return 1;
    if (t <= 2)
        return 2;
    if (t <= 4)
        return 4;
    if (t <= 8)
        return 8;
    return 12;
}

// Wide products use 8 waves (256 rows) per workgroup so every X chunk staged
// in LDS serves twice the rows; narrow ones 4 waves (more workgroups).
#ifndef RAD_GEMM_PF
#define RAD_GEMM_PF 1
#endif
#ifndef RAD_GEMM_WVS
#define RAD_GEMM_WVS 12
#endif
// Rows of one workgroup of a product.
static uint32_t gemm_wg_rows(uint32_t rt, bool split)
{
    const uint32_t wv = rt >= 4 ? 8u : 4u;
    return (rt >= 12 && !split ? (uint32_t)RAD_GEMM_WVS : wv) * 32u;
}

GemmSplit gemm_split(uint32_t rt, bool split, uint32_t npad, uint32_t S, uint32_t ncu)
{
    const uint32_t rows = gemm_wg_rows(rt, split);
    const uint32_t tiles = (npad + rows - 1) / rows, nch = npad >> 6;
    // row tiles whose tiles x S workgroups make whole rounds over the CUs
    // keep S slices; the rest, fewer than ncu / S tiles, get S2 slices
    const uint32_t full = tiles * S / ncu * ncu / S;
    if (full >= tiles || ncu % S != 0)
        return GemmSplit{S, S, tiles * rows};
    uint32_t S2 = ncu / (tiles - full);
    if (S2 > nch)
        S2 = nch;
    if (S2 < S)
        S2 = S;
    return GemmSplit{S, S2, full * rows};
}

template <int RT>
static void gemm_rt(bool split, const uint16_t *fhi, const uint16_t *flo, const uint16_t *xhi,
                    const uint16_t *xlo, uint32_t npad, GemmSplit ks, float *yp, hipStream_t st)
{
    constexpr int WV = RT >= 4 ? 8 : 4;
    // the widest single-half sweep: 12 wavefronts (3 per SIMD at 154
    // VGPRs), so each CU has 12 chunks of A in flight instead of 8
    constexpr int WVS = RT >= 12 ? RAD_GEMM_WVS : WV;
    // two chunks of A in flight for the single-half sweep; the split product
    // already holds both halves of A and X (PF = 2 would cost occupancy)
    constexpr int PFS = RT >= 4 ? RAD_GEMM_PF : 1;
    const uint32_t rows = gemm_wg_rows(RT, split);
    const uint32_t tiles = (npad + rows - 1) / rows, head = ks.rsplit / rows;
    const uint32_t nwg = head * ks.S + (tiles - head) * ks.S2;
    if (split)
        k_gemm<RT, true, WV, 1><<<nwg, WV * 64, 0, st>>>(fhi, flo, xhi, xlo, npad, ks, yp);
    else
        k_gemm<RT, false, WVS, PFS><<<nwg, WVS * 64, 0, st>>>(fhi, flo, xhi, xlo, npad, ks, yp);
}

void launch_gemm(uint32_t rt, bool split, const uint16_t *fhi, const uint16_t *flo,
                 const uint16_t *xhi, const uint16_t *xlo, uint32_t npad, GemmSplit ks,
                 float *yp, hipStream_t st)
{
    switch (rt) {
    case 1: gemm_rt<1>(split, fhi, flo, xhi, xlo, npad, ks, yp, st); break;
    case 2: gemm_rt<2>(split, fhi, flo, xhi, xlo, npad, ks, yp, st); break;
    case 4: gemm_rt<4>(split, fhi, flo, xhi, xlo, npad, ks, yp, st); break;
    case 8: gemm_rt<8>(split, fhi, flo, xhi, xlo, npad, ks, yp, st); break;
    default: gemm_rt<12>(split, fhi, flo, xhi, xlo, npad, ks, yp, st); break;
    }
}

void launch_gather_update(int kind, const float *yp, GemmSplit ks, uint32_t rp, uint32_t R,
                          uint32_t npad, uint32_t n, const float *E, const float *refl,
                          const float *area, float *rv, float *B, uint16_t *xhi, uint16_t *xlo,
                          double *parts, uint32_t nblocks, hipStream_t st)
{
    k_gather_update<<<nblocks, 256, 0, st>>>(kind, yp, ks, rp, R, npad, n, E, refl, area, rv, B,
                                             xhi, xlo, parts);
}

void launch_gather_out(uint32_t R, uint32_t npad, uint32_t n, const float *B, const float *rv,
                       float *out, hipStream_t st)
{
    k_gather_out<<<1024, 256, 0, st>>>(R, npad, n, B, rv, out);
}

void launch_pad_columns(uint32_t R, uint32_t npad, uint32_t n, const float *src, float *dst,
                        uint16_t *xhi, uint16_t *xlo, hipStream_t st)
{
    k_pad_columns<<<1024, 256, 0, st>>>(R, npad, n, src, dst, xhi, xlo);
}

void launch_sum_partials(uint32_t R, uint32_t npad, uint32_t n, GemmSplit ks, uint32_t rp,
                         const float *yp, float *out, hipStream_t st)
{
    k_sum_partials<<<1024, 256, 0, st>>>(R, npad, n, ks, rp, yp, out);
}

} // namespace rad

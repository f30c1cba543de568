// rad_pack.hip - many independent scenes in one launch.
//
// The scenes of a pack are concatenated (triangles, patches) with per-scene
// offsets.  Each pass gives every scene that has not reached the residual
// target its own shooter slots (the POWER rule inside the scene); the
// rasterizer kernels draw a slot's hemicubes against its scene's triangle
// range only and reduce into F rows indexed from the scene's first patch;
// the shoot kernel applies a scene's slots to that scene's receivers.  One
// pass is therefore one launch of each kernel for all scenes together.
#include "rad_internal.h"

namespace rad {

// Per-scene selection: one workgroup per scene.  FP64 residual
// sum_A (r+g+b)(unshot) / den; a scene at or below the target is done (the
// pass it finished in is kept); otherwise its k shooters are the top-k by
// A_i (r+g+b) with ties to the lower index, found by k rounds of a block
// max over packed keys (power bits << 32 | ~index), each round below the
// previous key.  Zero-power patches are never chosen.
__global__ __launch_bounds__(256) void k_pack_select(
    const uint32_t *__restrict__ sc_pbeg, const uint32_t *__restrict__ sc_slots,
    const double *__restrict__ sc_den, const float *__restrict__ unshot,
    const float *__restrict__ area, const uint32_t *__restrict__ pass, double target,
    uint32_t kmax, uint32_t *__restrict__ sc_done, double *__restrict__ sc_res,
    uint32_t *__restrict__ sc_nsel, uint32_t *__restrict__ sc_shot, uint32_t *__restrict__ cand)
{
    __shared__ double dred[4];
    __shared__ uint64_t kred[4];
    const uint32_t sc = blockIdx.x;
    const uint32_t pb = sc_pbeg[sc], pe = sc_pbeg[sc + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double num = 0.0;
    for (uint32_t j = pb + threadIdx.x; j < pe; j += blockDim.x) {
        const float *u = unshot + 3 * (size_t)j;
        num += (double)area[j] * ((double)u[0] + (double)u[1] + (double)u[2]);
    }
    for (int o = 32; o > 0; o >>= 1)
        num += __shfl_down(num, o, 64);
    if (lane == 0)
        dred[w] = num;
    __syncthreads();
    const double den = sc_den[sc];
    const double res = den > 0.0 ? (dred[0] + dred[1] + dred[2] + dred[3]) / den : 0.0;
    const bool done = sc_done[sc] != 0xFFFFFFFFu || res <= target;
    __syncthreads();
    if (done) {
        if (threadIdx.x == 0) {
            if (sc_done[sc] == 0xFFFFFFFFu)
                sc_done[sc] = *pass;
            sc_res[sc] = res;
            sc_nsel[sc] = 0;
        }
        return;
    }
    const uint32_t k = sc_slots[sc] < kmax ? sc_slots[sc] : kmax;
    uint64_t prev = ~0ull;
    uint32_t nsel = 0;
    for (uint32_t r = 0; r < k; ++r) {
        uint64_t best = 0;
        for (uint32_t j = pb + threadIdx.x; j < pe; j += blockDim.x) {
            const float *u = unshot + 3 * (size_t)j;
            float p = area[j] * (u[0] + u[1] + u[2]);
            if (!(p > 0.0f))
                continue;
            uint64_t key = ((uint64_t)__float_as_uint(p) << 32) | (uint64_t)(~j);
            if (key < prev && key > best)
                best = key;
        }
        for (int o = 32; o > 0; o >>= 1) {
            uint64_t v = __shfl_down(best, o, 64);
            best = v > best ? v : best;
        }
        if (lane == 0)
            kred[w] = best;
        __syncthreads();
        uint64_t b = kred[0];
        for (int q = 1; q < 4; ++q)
            b = kred[q] > b ? kred[q] : b;
        __syncthreads();
        if (b == 0)
            break;
        if (threadIdx.x == 0)
            cand[(size_t)sc * kmax + r] = ~(uint32_t)(b & 0xFFFFFFFFull);
        prev = b;
        ++nsel;
    }
    if (threadIdx.x == 0) {
        sc_res[sc] = res;
        sc_nsel[sc] = nsel; // the slot kernel may give the scene fewer
    }
}

// Slots of the pass, one workgroup.  Each active scene found sc_nsel <=
// its quota of shooters; when they all fit in the K slots every scene keeps
// them, otherwise the slots are shared out by water-filling: every scene
// gets min(nsel, L) for the largest level L whose total fits, and the slots
// left over go one each to scenes that wanted more, starting at a scene that
// moves on every pass (so none is passed over for long).  sc_nsel becomes
// the number the scene shoots, sc_shot counts them.  Slots are laid out
// scene after scene; the slot arrays carry the scene's triangle range, first
// patch, and, for scenes at their own resolution, hemicube side and ΔFF
// table offset.
__device__ static inline uint32_t block_sum_u32(uint32_t v, uint32_t *red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w)
        t += red[w];
    return t;
}

// Inclusive scan of one value per thread through part[] (blockDim.x <= 1024).
__device__ static inline uint32_t block_scan_u32(uint32_t v, uint32_t *part)
{
    __syncthreads();
    part[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        uint32_t o = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += o;
        __syncthreads();
    }
    return part[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_pack_slots(
    uint32_t nscene, const uint32_t *__restrict__ sc_tbeg, const uint32_t *__restrict__ sc_pbeg,
    uint32_t *__restrict__ sc_nsel, uint32_t *__restrict__ sc_shot,
    const uint32_t *__restrict__ cand, uint32_t kmax, uint32_t K,
    uint32_t *__restrict__ sc_slot_beg, uint32_t *__restrict__ shooters,
    uint32_t *__restrict__ slot_tbeg, uint32_t *__restrict__ slot_tcnt,
    uint32_t *__restrict__ slot_pbeg, const uint32_t *__restrict__ sc_w,
    const uint32_t *__restrict__ sc_dffo, uint32_t *__restrict__ slot_w,
    uint32_t *__restrict__ slot_dffo, uint32_t *__restrict__ nshoot, uint32_t *__restrict__ pass)
{
    __shared__ uint32_t part[1024];
    __shared__ uint32_t red[16];
    // each thread owns a contiguous run of scenes
    const uint32_t per = (nscene + blockDim.x - 1) / blockDim.x;
    const uint32_t a = threadIdx.x * per < nscene ? threadIdx.x * per : nscene;
    const uint32_t b = a + per < nscene ? a + per : nscene;
    auto level_total = [&](uint32_t L) {
        uint32_t t = 0;
        for (uint32_t s = a; s < b; ++s)
            t += min(sc_nsel[s], L);
        return block_sum_u32(t, red);
    };
    // largest L in [0, kmax] with sum_s min(nsel_s, L) <= K (binary search;
    // the total only grows with L)
    uint32_t lo = 0, hi = kmax;
    if (level_total(kmax) <= K) {
        lo = kmax;
    } else {
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (level_total(mid) <= K)
                lo = mid;
            else
                hi = mid;
        }
    }
    const uint32_t L = lo;
    const uint32_t left = K - level_total(L);
    // the scenes wanting more than L, ranked from a start scene that moves
    // on every pass; the first `left` of them get one slot more
    const uint32_t start = nscene ? *pass % nscene : 0u;
    uint32_t want_mine = 0, r0_mine = 0;
    for (uint32_t s = a; s < b; ++s) {
        want_mine += sc_nsel[s] > L ? 1u : 0u;
        r0_mine += (s < start && sc_nsel[s] > L) ? 1u : 0u;
    }
    const uint32_t want_before = block_scan_u32(want_mine, part) - want_mine;
    const uint32_t nwant = part[blockDim.x - 1];
    const uint32_t r0 = block_sum_u32(r0_mine, red);
    auto count_of = [&](uint32_t s, uint32_t &rank) {
        uint32_t n = min(sc_nsel[s], L);
        if (sc_nsel[s] > L) {
            const uint32_t rot = (rank + nwant - r0) % nwant;
            n += rot < left ? 1u : 0u;
            ++rank;
        }
        return n;
    };
    uint32_t mine = 0;
    {
        uint32_t rank = want_before;
        for (uint32_t s = a; s < b; ++s)
            mine += count_of(s, rank);
    }
    const uint32_t incl = block_scan_u32(mine, part);
    const uint32_t total = part[blockDim.x - 1];
    uint32_t off = incl - mine, rank = want_before;
    for (uint32_t s = a; s < b; ++s) {
        const uint32_t n = count_of(s, rank);
        sc_slot_beg[s] = off;
        sc_nsel[s] = n;
        sc_shot[s] += n;
        for (uint32_t r = 0; r < n; ++r) {
            const uint32_t slot = off + r;
            if (slot < K) {
                shooters[slot] = cand[(size_t)s * kmax + r];
                slot_tbeg[slot] = sc_tbeg[s];
                slot_tcnt[slot] = sc_tbeg[s + 1] - sc_tbeg[s];
                slot_pbeg[slot] = sc_pbeg[s];
                if (sc_w) {
                    slot_w[slot] = sc_w[s];
                    slot_dffo[slot] = sc_dffo[s];
                }
            }
        }
        off += n;
    }
    if (threadIdx.x == 0) {
        *nshoot = total < K ? total : K;
        *pass += 1u;
    }
}

// k_shoot for a pack: receiver j takes the slots of its own scene, F row
// entry j - first patch of the scene (rows are rowlen long).  Like k_shoot
// it sets every nonzero F entry it reads back to zero.
__global__ __launch_bounds__(256) void k_pack_shoot(
    uint32_t mode, const uint32_t *__restrict__ shooters, const uint32_t *__restrict__ nshoot,
    const float *__restrict__ snap, float *__restrict__ ff, uint32_t rowlen,
    const uint32_t *__restrict__ patch_scene, const uint32_t *__restrict__ sc_pbeg,
    const uint32_t *__restrict__ sc_slot_beg, const uint32_t *__restrict__ sc_nsel,
    const float *__restrict__ refl, const float *__restrict__ area, uint32_t npatch,
    float *__restrict__ unshot, float *__restrict__ shot)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npatch)
        return;
    const uint32_t sc = patch_scene[j];
    const uint32_t n = sc_nsel[sc];
    if (n == 0)
        return;
    const uint32_t s0 = sc_slot_beg[sc], jl = j - sc_pbeg[sc], ns = *nshoot;
    float u0 = unshot[3 * (size_t)j + 0], u1 = unshot[3 * (size_t)j + 1],
          u2 = unshot[3 * (size_t)j + 2];
    float r0 = refl[3 * (size_t)j + 0], r1 = refl[3 * (size_t)j + 1], r2 = refl[3 * (size_t)j + 2];
    float inv_aj = 1.0f / area[j];
    for (uint32_t s = s0; s < s0 + n && s < ns; ++s) {
        float f = ff[(size_t)s * rowlen + jl];
        if (f == 0.0f)
            continue;
        ff[(size_t)s * rowlen + jl] = 0.0f; // F rows left zero for the next pass
        uint32_t i = shooters[s];
        float d0 = snap[3 * s + 0], d1 = snap[3 * s + 1], d2 = snap[3 * s + 2];
        if (mode == RAD_MODE_REFERENCE) {
            u0 += d0 * f * 0.3f * refl[3 * (size_t)i + 0];
            u1 += d1 * f * 0.3f * refl[3 * (size_t)i + 1];
            u2 += d2 * f * 0.3f * refl[3 * (size_t)i + 2];
        } else {
            float g = f * (area[i] * inv_aj);
            u0 += r0 * d0 * g;
            u1 += r1 * d1 * g;
            u2 += r2 * d2 * g;
        }
    }
    for (uint32_t s = s0; s < s0 + n && s < ns; ++s) {
        if (shooters[s] == j) {
            float d0 = snap[3 * s + 0], d1 = snap[3 * s + 1], d2 = snap[3 * s + 2];
            shot[3 * (size_t)j + 0] += d0;
            shot[3 * (size_t)j + 1] += d1;
            shot[3 * (size_t)j + 2] += d2;
            u0 -= d0;
            u1 -= d1;
            u2 -= d2;
        }
    }
    unshot[3 * (size_t)j + 0] = u0;
    unshot[3 * (size_t)j + 1] = u1;
    unshot[3 * (size_t)j + 2] = u2;
}

void launch_pack_select(const PackBuffers &pk, const float *unshot, const float *area,
                        hipStream_t st)
{
    k_pack_select<<<pk.nscene, 256, 0, st>>>(pk.sc_pbeg, pk.sc_slots, pk.sc_den, unshot, area,
                                             pk.pass, pk.target, pk.kmax, pk.sc_done, pk.sc_res,
                                             pk.sc_nsel, pk.sc_shot, pk.cand);
    k_pack_slots<<<1, 1024, 0, st>>>(pk.nscene, pk.sc_tbeg, pk.sc_pbeg, pk.sc_nsel, pk.sc_shot,
                                     pk.cand, pk.kmax, pk.K, pk.sc_slot_beg, pk.shooters,
                                     pk.slot_tbeg, pk.slot_tcnt, pk.slot_pbeg, pk.sc_w, pk.sc_dffo,
                                     pk.slot_w, pk.slot_dffo, pk.nshoot, pk.pass);
}

void launch_pack_shoot(const PackBuffers &pk, uint32_t mode, const float *snap, float *ff,
                       uint32_t rowlen, const float *refl, const float *area, uint32_t npatch,
                       float *unshot, float *shot, hipStream_t st)
{
    k_pack_shoot<<<(npatch + 255) / 256, 256, 0, st>>>(
        mode, pk.shooters, pk.nshoot, snap, ff, rowlen, pk.patch_scene, pk.sc_pbeg,
        pk.sc_slot_beg, pk.sc_nsel, refl, area, npatch, unshot, shot);
}

} // namespace rad

// k_components.hip -- connectedComponentsWithStats on the GPU
// (src/imgproc/components.zig:408-535, C ABI capi.zig:533-550).
//
// The labels equal the reference's. With 4-connectivity, components are
// numbered in raster order of their first pixel. With 8-connectivity they
// are numbered by their first 2x2 block in block-raster order: the block row
// of their top row, then the smallest block column they touch in that block
// row (components.zig:273-328). Two 8-components never share a 2x2 block, so
// these keys are distinct.
//
//   cc_init      L[i] = i for nonzero pixels, the "none" value otherwise.
//   cc_merge     joins each nonzero pixel to its earlier neighbours (left,
//                up; up-left and up-right for 8-connectivity). The union-find
//                keeps L[x] <= x: a root is attached under the smaller root
//                with atomicMin, so every root is the component's first
//                pixel in raster order.
//   cc_flatten   points every pixel at its root; for 8-connectivity, takes
//                the minimum block column over the root's block row.
//   cc_mark      marks each component's key; an exclusive scan of the marks
//                is the component's rank, its label minus one.
//   cc_root_lab  the final label of each root.
//   cc_runs      writes the label image and, per horizontal run, adds the
//                run to its label's integer statistics (area, bounding box,
//                twice the x sum, the y sum), the sums of components.zig:178-210.
//   cc_rows      per row: nonzero count and x sum, first and last zero, from
//                which the host forms the background's statistics.
#include "components_gpu.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kScanBlock = 1024;  // marks per scan block (256 threads x 4)

// Parents only decrease and stay inside the component, so a stale read is
// still a valid ancestor; cc_union rechecks with atomicMin.
__device__ __attribute__((always_inline)) inline uint32_t cc_find(const uint32_t *L, uint32_t x) {
    uint32_t p = __atomic_load_n(L + x, __ATOMIC_RELAXED);
    while (p != x) {
        x = p;
        p = __atomic_load_n(L + x, __ATOMIC_RELAXED);
    }
    return x;
}

__device__ __attribute__((always_inline)) inline void cc_union(uint32_t *L, uint32_t a, uint32_t b) {
    for (;;) {
        a = cc_find(L, a);
        b = cc_find(L, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        // a > b: attach root a under b, unless another thread moved a first.
        const uint32_t old = atomicMin(L + a, b);
        if (old == a) return;
        a = old;
    }
}

__global__ __launch_bounds__(256) void cc_init(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    const uint32_t y = i / c.w, x = i - y * c.w;
    c.L[i] = c.src[(size_t)y * c.sstride + x] ? i : kNone;
    c.key[i] = kNone;
}

__global__ __launch_bounds__(256) void cc_merge(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    const uint32_t y = i / c.w, x = i - y * c.w;
    const uint8_t *row = c.src + (size_t)y * c.sstride;
    if (!row[x]) return;
    const bool left = x > 0 && row[x - 1];
    if (left) cc_union(c.L, i, i - 1);
    if (y == 0) return;
    const uint8_t *up = row - c.sstride;
    const bool upc = up[x] != 0;
    if (upc) cc_union(c.L, i, i - c.w);
    if (!c.eight) return;
    // The diagonals are already joined through a nonzero left or up pixel.
    if (x > 0 && !left && !upc && up[x - 1]) cc_union(c.L, i, i - c.w - 1);
    if (x + 1 < c.w && !upc && up[x + 1]) cc_union(c.L, i, i - c.w + 1);
}

__global__ __launch_bounds__(256) void cc_flatten(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    if (c.L[i] == kNone) return;
    const uint32_t r = cc_find(c.L, i);
    c.L[i] = r;
    if (c.eight) {
        const uint32_t y = i / c.w, x = i - y * c.w, ry = r / c.w;
        if ((y >> 1) == (ry >> 1)) atomicMin(c.key + r, x >> 1);
    }
}

// The key of root r: its raster index, or its block row and the smallest
// block column of the component in that row.
__device__ __attribute__((always_inline)) inline uint32_t cc_key(const CcArgs &c, uint32_t r) {
    if (!c.eight) return r;
    return (r / c.w >> 1) * c.kw + c.key[r];
}

__global__ __launch_bounds__(256) void cc_clear(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(256) void cc_mark(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    if (c.L[i] == i) c.mark[cc_key(c, i)] = 1;
}

// Exclusive scan of the marks in place: block sums, their scan, the adds.
__device__ __attribute__((always_inline)) inline uint32_t block_exclusive(uint32_t v, uint32_t *s, uint32_t *total) {
    // Inclusive scan within the wave, then across the 4 waves.
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    if (lane == 63) s[wv] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < wv; k++) before += s[k];
    if (total) *total = s[0] + s[1] + s[2] + s[3];
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(256) void cc_scan_blocks(uint32_t *mark, uint32_t nkeys, uint32_t *bsum) {
    __shared__ uint32_t s[4];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < nkeys ? mark[base + k] : 0;
        t += v[k];
    }
    uint32_t total;
    uint32_t e = block_exclusive(t, s, &total);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < nkeys) mark[base + k] = e;
        e += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the block sums in place; bsum[nb] = total.
__global__ __launch_bounds__(256) void cc_scan_sums(uint32_t *bsum, uint32_t nb) {
    __shared__ uint32_t s[4];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0;
        uint32_t total;
        const uint32_t e = block_exclusive(v, s, &total);
        if (i < nb) bsum[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(256) void cc_root_lab(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    if (c.L[i] != i) return;
    const uint32_t k = cc_key(c, i);
    c.lab[i] = c.bsum[k / kScanBlock] + c.mark[k] + 1;
}

__global__ __launch_bounds__(256) void cc_runs(CcArgs c) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= c.n) return;
    const uint32_t y = i / c.w, x = i - y * c.w;
    const uint32_t r = c.L[i];
    const uint32_t l = r == kNone ? 0u : c.lab[r];
    if (c.labels) *(int32_t *)((uint8_t *)c.labels + (size_t)y * c.lstride + 4u * x) = (int32_t)l;
    if (r == kNone) return;
    if (x > 0 && c.L[i - 1] != kNone) return;  // not the start of a run
    uint32_t x1 = x;
    while (x1 + 1 < c.w && c.L[i + (x1 + 1 - x)] != kNone) x1++;
    const uint64_t len = x1 - x + 1;
    uint64_t *a = c.acc + (size_t)l * 4;
    uint32_t *b = c.box + (size_t)l * 4;
    atomicAdd((unsigned long long *)a, (unsigned long long)len);
    atomicAdd((unsigned long long *)(a + 1), (unsigned long long)(len * ((uint64_t)x + x1)));
    atomicAdd((unsigned long long *)(a + 2), (unsigned long long)(len * y));
    atomicMin(b, x);
    atomicMin(b + 1, y);
    atomicMax(b + 2, x1);
    atomicMax(b + 3, y);
}

__global__ __launch_bounds__(256) void cc_clear_stats(CcArgs c, uint32_t nlab) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= nlab) return;
    uint64_t *a = c.acc + (size_t)i * 4;
    a[0] = a[1] = a[2] = a[3] = 0;
    uint32_t *b = c.box + (size_t)i * 4;
    b[0] = b[1] = kNone;
    b[2] = b[3] = 0;
}

// One wave per row: zero count, nonzero x sum (twice), first and last zero.
__global__ __launch_bounds__(256) void cc_rows(CcArgs c) {
    const uint32_t y = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (y >= c.h) return;
    const int lane = threadIdx.x & 63;
    const uint8_t *row = c.src + (size_t)y * c.sstride;
    uint32_t zeros = 0, first = kNone, last = 0;
    uint64_t fgx2 = 0;
    for (uint32_t x0 = 0; x0 < c.w; x0 += 64) {
        const uint32_t x = x0 + lane;
        const bool in = x < c.w;
        const bool fg = in && row[x] != 0;
        const uint64_t zb = __ballot(in && !fg);
        if (fg) fgx2 += 2ull * x;
        if (zb) {
            zeros += __popcll(zb);
            first = min(first, x0 + (uint32_t)__ffsll((long long)zb) - 1u);
            last = max(last, x0 + 63u - (uint32_t)__clzll((long long)zb));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fgx2 += __shfl_xor(fgx2, o);
    if (lane == 0) {
        c.rows[(size_t)y * 4] = zeros;
        c.rows[(size_t)y * 4 + 1] = first;
        c.rows[(size_t)y * 4 + 2] = last;
        c.rows[(size_t)y * 4 + 3] = fgx2;
    }
}

inline unsigned blocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

size_t cc_scratch_bytes(uint32_t w, uint32_t h) {
    const uint64_t n = (uint64_t)w * h;
    const uint64_t nkeys = n;  // the 8-connected key space is smaller
    const uint64_t nb = (nkeys + kScanBlock - 1) / kScanBlock;
    const uint64_t nlab = n + 1;
    return (size_t)(4 * n * 3 + 4 * (nb + 1) + 8 * 4 * nlab + 4 * 4 * nlab + 8 * 4 * (uint64_t)h + 8 * 256);
}

int cc_label(CcArgs &c, void *scratch, hipStream_t st, uint32_t *count) {
    const uint64_t n = (uint64_t)c.w * c.h;
    c.n = (uint32_t)n;
    c.kw = (c.w + 1) / 2;
    const uint32_t nkeys = c.eight ? c.kw * ((c.h + 1) / 2) : c.n;
    const uint32_t nb = (nkeys + kScanBlock - 1) / kScanBlock;
    uint8_t *p = static_cast<uint8_t *>(scratch);
    auto take = [&](size_t bytes) {
        uint8_t *q = p;
        p += (bytes + 255) & ~(size_t)255;
        return q;
    };
    c.L = (uint32_t *)take(4 * n);
    c.key = (uint32_t *)take(4 * n);
    c.lab = c.key;  // the keys of 8-connectivity are read before the labels overwrite them
    c.mark = (uint32_t *)take(4 * (size_t)nkeys);
    c.bsum = (uint32_t *)take(4 * ((size_t)nb + 1));
    c.acc = (uint64_t *)take(8 * 4 * (n + 1));
    c.box = (uint32_t *)take(4 * 4 * (n + 1));
    c.rows = (uint64_t *)take(8 * 4 * (size_t)c.h);
    hipLaunchKernelGGL(cc_init, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_merge, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_flatten, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_clear, dim3(blocks(nkeys)), dim3(256), 0, st, c.mark, nkeys);
    hipLaunchKernelGGL(cc_mark, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_scan_blocks, dim3(nb), dim3(256), 0, st, c.mark, nkeys, c.bsum);
    hipLaunchKernelGGL(cc_scan_sums, dim3(1), dim3(256), 0, st, c.bsum, nb);
    if (hipMemcpyAsync(c.host_total, c.bsum + nb, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return INTERNAL;
    const uint32_t total = *c.host_total;
    const uint32_t nlab = total + 1;
    hipLaunchKernelGGL(cc_root_lab, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_clear_stats, dim3(blocks(nlab)), dim3(256), 0, st, c, nlab);
    hipLaunchKernelGGL(cc_runs, dim3(blocks(n)), dim3(256), 0, st, c);
    hipLaunchKernelGGL(cc_rows, dim3((c.h + 3) / 4), dim3(256), 0, st, c);
    *count = nlab;
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

}  // namespace visar

// 26-connected component labelling of a cell mask on the device: the
// island test of the SIMP analysis, scipy.ndimage.label(g, structure =
// ones((3, 3, 3))) as simp._Analysis._connected and simp._prune_dead call it
// (simp.py:264-282, 1276-1282), with the reference's numbering: components
// 1..n in the order of their first cell in C order.
//
// Each foreground cell starts as its own root (its linear index). Rounds of
// hooking (a cell's root is hooked to the smallest root among its 26
// neighbours with an atomic min on the root) and pointer jumping (every
// cell follows its root's root to the end) repeat until a round changes
// nothing; the root of a component is then its smallest cell index, so the
// result does not depend on the order the rounds ran in. Numbering: a
// prefix count of the roots in C order.
#include <mutex>
#include <string>
#include <vector>

#include "lp_internal.h"

namespace lp {
namespace {

constexpr int kB = 256;
constexpr int kNoAllocMaxCells = LP_LABEL26_NOALLOC_MAX_CELLS;
constexpr int kNoAllocBlocks = kNoAllocMaxCells / kB;

}  // namespace

// lp_label26_noalloc's working memory: [0] the round's changed flag, [1]
// the component count, [2 ..] one root count per block of kB cells, then
// (in place) their exclusive prefix. Host code reaches it only through the
// per-device address from hipGetSymbolAddress.
static __device__ __align__(256) int scratch[2 + kNoAllocBlocks];

namespace {

__global__ void k_label_init(const uint8_t* __restrict__ m, int n, int* lab) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) lab[i] = m[i] ? i : -1;
}

// the root of i (lab[r] == r), following parents
__device__ __forceinline__ int find_root(const int* lab, int i) {
    int r = lab[i];
    while (r >= 0 && lab[r] != r) r = lab[r];
    return r;
}

// hook: the root of each foreground cell takes the smallest root among the
// cell's foreground neighbours (atomic min: roots only decrease)
__global__ void k_label_hook(int nx, int ny, int nz, int* lab, int* changed) {
    const int n = nx * ny * nz;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || lab[i] < 0) return;
    const int x = i / (ny * nz), y = (i / nz) % ny, z = i % nz;
    const int ri = find_root(lab, i);
    int best = ri;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int X = x + dx, Y = y + dy, Z = z + dz;
                if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
                const int j = (X * ny + Y) * nz + Z;
                if (lab[j] < 0) continue;
                const int rj = find_root(lab, j);
                best = rj < best ? rj : best;
            }
    if (best < ri) {
        atomicMin(&lab[ri], best);
        *changed = 1;
    }
}

// pointer jumping: every cell points at its root
__global__ void k_label_jump(int n, int* lab) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || lab[i] < 0) return;
    lab[i] = find_root(lab, i);
}

// roots in C order -> 1..n: flag[i] = 1 where i is a root
__global__ void k_label_roots(int n, const int* lab, int* flag) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) flag[i] = lab[i] == i ? 1 : 0;
}

__global__ void k_label_number(int n, const int* lab, const int* num, int* out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) out[i] = lab[i] < 0 ? 0 : num[lab[i]];
}

// ---- the allocation-free variant: parents in the caller's label array,
// the round flag and the numbering's block counts in a static array ------

__global__ void k_flag_clear(int* flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 0;
}

// per block: the roots among its cells (lab[i] == i)
__global__ void k_root_counts(int n, const int* __restrict__ lab, int* __restrict__ counts) {
    __shared__ int warp_sums[kB / 64];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int root = (i < n && lab[i] == i) ? 1 : 0;
    const unsigned long long bal = __ballot(root);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) warp_sums[w] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < kB / 64; ++k) t += warp_sums[k];
        counts[blockIdx.x] = t;
    }
}

// one block: exclusive prefix of nb counts in place, the total to *total
__global__ void k_scan_counts(int nb, int* counts, int* total) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int a = t * per, b = min(a + per, nb);
    int s = 0;
    for (int k = a; k < b; ++k) s += counts[k];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int k = 0; k < 1024; ++k) {
            const int v = part[k];
            part[k] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    int run = part[t];
    for (int k = a; k < b; ++k) {
        const int v = counts[k];
        counts[k] = run;
        run += v;
    }
}

// each root gets its number (1 + the roots before it in C order), stored
// encoded as -(number) - 1 <= -2 in its own slot (background stays -1)
__global__ void k_number_roots(int n, int* lab, const int* __restrict__ offsets) {
    __shared__ int warp_sums[kB / 64];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int root = (i < n && lab[i] == i) ? 1 : 0;
    const unsigned long long bal = __ballot(root);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) warp_sums[w] = __popcll(bal);
    __syncthreads();
    int before = offsets[blockIdx.x];
    for (int k = 0; k < w; ++k) before += warp_sums[k];
    before += __popcll(bal & ((1ull << lane) - 1ull));
    if (root) lab[i] = -(before + 1) - 1;
}

// cells that are not roots read their root's number (root slots are only
// rewritten by the next kernel, so no cell reads a rewritten slot)
__global__ void k_decode_members(int n, int* lab) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n && lab[i] >= 0) lab[i] = -lab[lab[i]] - 1;
}

__global__ void k_decode_roots(int n, int* lab) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int v = lab[i];
    if (v == -1) lab[i] = 0;
    else if (v <= -2) lab[i] = -v - 1;
}

}  // namespace
}  // namespace lp

namespace lp {

// components of a device cell mask on stream s: labels = each cell's root
// (its component's smallest cell index), -1 off the mask; flag is a device
// int for the rounds' changed bit (the host reads it after each round)
void label26_roots(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* flag, hipStream_t s) {
    const int n = nx * ny * nz;
    const unsigned nb = (unsigned)((n + kB - 1) / kB);
    hipLaunchKernelGGL(k_label_init, dim3(nb), dim3(kB), 0, s, mask, n, labels);
    LP_HIP(hipGetLastError());
    for (int round = 0; round < n + 1; ++round) {
        hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, s, flag);
        hipLaunchKernelGGL(k_label_hook, dim3(nb), dim3(kB), 0, s, nx, ny, nz, labels, flag);
        hipLaunchKernelGGL(k_label_jump, dim3(nb), dim3(kB), 0, s, n, labels);
        LP_HIP(hipGetLastError());
        int changed = 0;
        LP_HIP(hipMemcpyAsync(&changed, flag, 4, hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        if (!changed) break;
    }
}

}  // namespace lp

using namespace lp;

extern "C" {

int lp_label26(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* count) {
    return api_call([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !mask || !labels || !count ||
            (size_t)nx * ny * nz >= (size_t)0x7fffffff) {
            set_error("lp_label26: bad argument");
            return LP_ERR_ARG;
        }
        const int n = nx * ny * nz;
        const unsigned nb = (unsigned)((n + kB - 1) / kB);
        DBuf dm, dl, dc, dout;
        dm.alloc(n);
        dl.alloc((size_t)n * 4);
        dc.alloc(4);
        dout.alloc((size_t)n * 4);
        LP_HIP(hipMemcpy(dm.p, mask, n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_label_init, dim3(nb), dim3(kB), 0, nullptr, dm.as<uint8_t>(), n, dl.as<int>());
        LP_HIP(hipGetLastError());
        // rounds until one changes nothing: each lowers at least one root
        // (roots only decrease, parents always point to smaller indices, so
        // there are no cycles), and n + 1 rounds bound them
        for (int round = 0; round < n + 1; ++round) {
            int changed = 0;
            LP_HIP(hipMemcpy(dc.p, &changed, 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_label_hook, dim3(nb), dim3(kB), 0, nullptr, nx, ny, nz, dl.as<int>(), dc.as<int>());
            hipLaunchKernelGGL(k_label_jump, dim3(nb), dim3(kB), 0, nullptr, n, dl.as<int>());
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpy(&changed, dc.p, 4, hipMemcpyDeviceToHost));
            if (!changed) break;
        }
        // number the roots in C order (host prefix sum over the root flags)
        hipLaunchKernelGGL(k_label_roots, dim3(nb), dim3(kB), 0, nullptr, n, dl.as<int>(), dout.as<int>());
        LP_HIP(hipGetLastError());
        std::vector<int> flag(n);
        LP_HIP(hipMemcpy(flag.data(), dout.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        int k = 0;
        for (int i = 0; i < n; ++i) flag[i] = flag[i] ? ++k : 0;
        LP_HIP(hipMemcpy(dout.p, flag.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        DBuf dnum;
        dnum.alloc((size_t)n * 4);
        hipLaunchKernelGGL(k_label_number, dim3(nb), dim3(kB), 0, nullptr, n, dl.as<int>(), dout.as<int>(),
                           dnum.as<int>());
        LP_HIP(hipGetLastError());
        LP_HIP(hipMemcpy(labels, dnum.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        *count = k;
        return LP_OK;
    });
}

}  // extern "C"

namespace {

// the device address of `scratch` on each GPU, taken once per device with
// that device current (each GPU holds its own copy of the array)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_dev[kMaxDevices] = {};

int* scratch_on_current_device() {
    int dev = 0;
    LP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("lp_label26_noalloc: device index out of range");
        throw ArgFailure();
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        LP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<int*>(p);
    }
    return g_scratch_dev[dev];
}

}  // namespace

extern "C" {

int lp_label26_noalloc(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* count) {
    return api_call([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !mask || !labels || !count) {
            set_error("lp_label26_noalloc: bad argument");
            return LP_ERR_ARG;
        }
        const size_t ncell = (size_t)nx * ny * nz;
        if (ncell > (size_t)kNoAllocMaxCells) {
            set_error("lp_label26_noalloc: " + std::to_string(ncell) + " cells exceed the limit of " +
                      std::to_string(kNoAllocMaxCells));
            return LP_ERR_ARG;
        }
        int* sc = scratch_on_current_device();
        int* flag = sc;
        int* total = sc + 1;
        int* counts = sc + 2;
        const int n = (int)ncell;
        const unsigned nb = (unsigned)((n + kB - 1) / kB);
        hipLaunchKernelGGL(k_label_init, dim3(nb), dim3(kB), 0, nullptr, mask, n, labels);
        LP_HIP(hipGetLastError());
        for (int round = 0; round < n + 1; ++round) {
            hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, nullptr, flag);
            hipLaunchKernelGGL(k_label_hook, dim3(nb), dim3(kB), 0, nullptr, nx, ny, nz, labels, flag);
            hipLaunchKernelGGL(k_label_jump, dim3(nb), dim3(kB), 0, nullptr, n, labels);
            LP_HIP(hipGetLastError());
            int changed = 0;
            LP_HIP(hipMemcpy(&changed, flag, 4, hipMemcpyDeviceToHost));
            if (!changed) break;
        }
        hipLaunchKernelGGL(k_root_counts, dim3(nb), dim3(kB), 0, nullptr, n, labels, counts);
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, nullptr, (int)nb, counts, total);
        hipLaunchKernelGGL(k_number_roots, dim3(nb), dim3(kB), 0, nullptr, n, labels, counts);
        hipLaunchKernelGGL(k_decode_members, dim3(nb), dim3(kB), 0, nullptr, n, labels);
        hipLaunchKernelGGL(k_decode_roots, dim3(nb), dim3(kB), 0, nullptr, n, labels);
        LP_HIP(hipGetLastError());
        int k = 0;
        LP_HIP(hipMemcpy(&k, total, 4, hipMemcpyDeviceToHost));
        *count = k;
        return LP_OK;
    });
}

}  // extern "C"

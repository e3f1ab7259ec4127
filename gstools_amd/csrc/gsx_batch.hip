// Many realizations of the randomisation-method sum in one persistent launch.
//
// Each item is one RandMeth / IncomprRandMeth call (src/gstools/field/
// generator.py:243-270, :529-567) with its own modes (cov, z1, z2), its own
// points (a (d, P) array or the 'ij' meshgrid of axes, tools/geometric.py:
// 340-356) and its own output.  Items of very different sizes (64^2 grids
// next to 1e6-point sets, N from 1000 to 8192) each underfill the GPU when
// launched alone; here they share one grid of resident blocks.
//
// Plan (host, once): items are cut into tiles of `tile` points; the tile
// list is ordered by descending N (longest-processing-time first) and copied
// to the device with the item descriptors.
//
// Launch: a one-thread kernel resets the tile counter (a fill kernel, so the
// sequence can be captured in a HIP graph), then a persistent kernel with
// (CUs x resident blocks) blocks pulls tiles with one atomicAdd per tile until
// the list is exhausted; every block exits when the counter passes the end.
// Tile work is the FP32 sum of gsx_fp32sum.h, the one the single launches run:
// FP64 phase in revolutions, magic-addition or fract reduction, v_cos_f32,
// blocked FP32 -> FP64 mode sums, modes staged through LDS in chunks of CHUNK
// (both PPT and CHUNK are autotuned variants).  Its segment and mode-group
// rules do not depend on PPT or CHUNK, so every variant, and the single
// launch, give a point the same bits.
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "gsx_fp32sum.h"
#include "gsx_internal.h"

namespace gsx {

struct BatchTile {
    int32_t item;
    int32_t pad;
    int64_t p0;
};

template <int D, bool INCOMPR, int PPT, int CHUNK, class Pts>
__device__ __forceinline__ void batch_tile(const gsx_batch_item& it, Pts pts, int64_t p0, double* tabd,
                                           float* tabf, double* s_wmax, int* s_nmagic) {
    constexpr int NOUT = INCOMPR ? D : 1;
    const int64_t P = it.P;
    double x[PPT][D];
    double acc[PPT][NOUT];
    double own = 0.0;
    const int64_t base = p0 + threadIdx.x;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
            pts.get(i, x[p]);
#pragma unroll
            for (int d = 0; d < D; ++d) own = coord_bound(own, x[p][d]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) x[p][d] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < NOUT; ++c) acc[p][c] = 0.0;
    }
    const double xseg = segment_xmax<D, PPT>(pts, P, p0, own, s_wmax);
    fp32_item_sum<D, PPT, INCOMPR, CHUNK>(it.N, it.cov, it.z1, it.z2, nullptr, xseg, x, acc, tabd, tabf,
                                          s_nmagic);
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
#pragma unroll
            for (int c = 0; c < NOUT; ++c) it.out[c * P + i] = (float)acc[p][c];
        }
    }
}

template <int D, bool INCOMPR, int PPT, int CHUNK>
__device__ __forceinline__ void batch_dispatch_pts(const gsx_batch_item& it, int64_t p0, double* tabd,
                                                   float* tabf, double* s_wmax, int* s_nmagic) {
    if (it.axes) {
        const double* a0 = it.axes;
        const double* a1 = D > 1 ? a0 + it.shape[0] : nullptr;
        const double* a2 = D > 2 ? a1 + it.shape[1] : nullptr;
        const uint32_t n1 = D > 1 ? (uint32_t)it.shape[1] : 1u;
        const uint32_t n2 = D > 2 ? (uint32_t)it.shape[2] : 1u;
        batch_tile<D, INCOMPR, PPT, CHUNK>(it, GridPoints<D>{a0, a1, a2, n1, n2}, p0, tabd, tabf, s_wmax,
                                           s_nmagic);
    } else {
        batch_tile<D, INCOMPR, PPT, CHUNK>(it, UnstructuredPoints<D>{it.pos, it.P}, p0, tabd, tabf, s_wmax,
                                           s_nmagic);
    }
}

__global__ void k_batch_reset(unsigned* __restrict__ counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

template <int PPT, int CHUNK>
__global__ __launch_bounds__(kThreads) void k_batch(const gsx_batch_item* __restrict__ items,
                                                    const BatchTile* __restrict__ tiles, int64_t ntiles,
                                                    unsigned* __restrict__ counter) {
    __shared__ double tabd[CHUNK * 4];
    __shared__ float tabf[CHUNK * 3];
    __shared__ double s_wmax[kWaves];
    __shared__ int s_nmagic[kWaves];
    __shared__ unsigned s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
        __syncthreads();
        const int64_t t = (int64_t)s_tile;
        __syncthreads();
        if (t >= ntiles) break;  // uniform across the block: every wave exits
        const BatchTile tl = tiles[t];
        const gsx_batch_item it = items[tl.item];
        if (it.kind == 1) {
            if (it.d == 2) batch_dispatch_pts<2, true, PPT, CHUNK>(it, tl.p0, tabd, tabf, s_wmax, s_nmagic);
            else batch_dispatch_pts<3, true, PPT, CHUNK>(it, tl.p0, tabd, tabf, s_wmax, s_nmagic);
        } else {
            if (it.d == 1) batch_dispatch_pts<1, false, PPT, CHUNK>(it, tl.p0, tabd, tabf, s_wmax, s_nmagic);
            else if (it.d == 2) batch_dispatch_pts<2, false, PPT, CHUNK>(it, tl.p0, tabd, tabf, s_wmax, s_nmagic);
            else batch_dispatch_pts<3, false, PPT, CHUNK>(it, tl.p0, tabd, tabf, s_wmax, s_nmagic);
        }
    }
}

typedef void (*BatchKernel)(const gsx_batch_item*, const BatchTile*, int64_t, unsigned*);

struct Variant {
    int ppt, chunk;
    BatchKernel fn;
};

static const Variant kVariants[] = {
    {2, 128, k_batch<2, 128>}, {2, 256, k_batch<2, 256>}, {4, 128, k_batch<4, 128>},
    {4, 256, k_batch<4, 256>}, {8, 128, k_batch<8, 128>}, {8, 256, k_batch<8, 256>},
};

}  // namespace gsx

struct gsx_batch_plan {
    gsx_batch_item* d_items = nullptr;
    gsx::BatchTile* d_tiles = nullptr;
    unsigned* d_counter = nullptr;
    int64_t ntiles = 0;
    gsx::Variant var{};
    int blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace gsx {

static int batch_enqueue(gsx_batch_plan* pl, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, pl->d_counter);
    int st = launch_status();
    if (st != GSX_OK || pl->ntiles == 0) return st;
    hipLaunchKernelGGL(pl->var.fn, dim3((unsigned)pl->blocks), dim3(kThreads), 0, s, pl->d_items, pl->d_tiles,
                       pl->ntiles, pl->d_counter);
    return launch_status();
}

static void plan_free(gsx_batch_plan* pl) {
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->d_items) (void)hipFree(pl->d_items);
    if (pl->d_tiles) (void)hipFree(pl->d_tiles);
    if (pl->d_counter) (void)hipFree(pl->d_counter);
    delete pl;
}

}  // namespace gsx

extern "C" {

int gsx_batch_plan_create(const gsx_batch_item* items, int64_t count, int64_t tile_points, int64_t mode_chunk,
                          gsx_batch_plan** plan) {
    using namespace gsx;
    if (!plan || count < 0 || (count > 0 && !items)) return GSX_E_INVALID;
    *plan = nullptr;
    const int64_t tp = tile_points > 0 ? tile_points : 1024;
    const int64_t mc = mode_chunk > 0 ? mode_chunk : 256;
    if ((tp != 512 && tp != 1024 && tp != 2048) || (mc != 128 && mc != 256)) return GSX_E_INVALID;
    const int ppt = (int)(tp / kThreads), chunk = (int)mc;
    Variant var{};
    for (const Variant& v : kVariants)
        if (v.ppt == ppt && v.chunk == chunk) var = v;
    if (!var.fn) return GSX_E_INVALID;

    std::vector<BatchTile> tiles;
    const int64_t per = (int64_t)kThreads * ppt;
    for (int64_t r = 0; r < count; ++r) {
        const gsx_batch_item& it = items[r];
        if (it.d < 1 || it.d > 3 || it.N < 0 || it.P < 0 || (it.kind != 0 && it.kind != 1)) return GSX_E_INVALID;
        if (it.kind == 1 && it.d < 2) return GSX_E_INVALID;
        if (it.P == 0) continue;
        if (!it.out || (it.N > 0 && (!it.cov || !it.z1 || !it.z2))) return GSX_E_INVALID;
        if (it.axes) {
            int64_t prod = 1;
            for (int i = 0; i < it.d; ++i) prod *= it.shape[i];
            if (prod != it.P || it.P >= ((int64_t)1 << 32)) return GSX_E_INVALID;
        } else if (!it.pos) {
            return GSX_E_INVALID;
        }
        for (int64_t p0 = 0; p0 < it.P; p0 += per) tiles.push_back(BatchTile{(int32_t)r, 0, p0});
    }
    // longest processing time first: cost of a tile ~ N (x d outputs)
    std::stable_sort(tiles.begin(), tiles.end(), [&](const BatchTile& a, const BatchTile& b) {
        const gsx_batch_item& ia = items[a.item];
        const gsx_batch_item& ib = items[b.item];
        return ia.N * (ia.kind ? ia.d : 1) > ib.N * (ib.kind ? ib.d : 1);
    });

    gsx_batch_plan* pl = new gsx_batch_plan();
    pl->var = var;
    pl->ntiles = (int64_t)tiles.size();
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(var.fn), kThreads, 0) !=
            hipSuccess) {
        plan_free(pl);
        return GSX_E_NO_DEVICE;
    }
    pl->blocks = std::max(1, cus * std::max(1, occ));
    bool ok = hipMalloc(&pl->d_counter, sizeof(unsigned)) == hipSuccess;
    if (ok && count > 0) ok = hipMalloc(&pl->d_items, sizeof(gsx_batch_item) * count) == hipSuccess;
    if (ok && pl->ntiles > 0) ok = hipMalloc(&pl->d_tiles, sizeof(BatchTile) * pl->ntiles) == hipSuccess;
    if (ok && count > 0)
        ok = hipMemcpy(pl->d_items, items, sizeof(gsx_batch_item) * count, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && pl->ntiles > 0)
        ok = hipMemcpy(pl->d_tiles, tiles.data(), sizeof(BatchTile) * pl->ntiles, hipMemcpyHostToDevice) ==
             hipSuccess;
    if (!ok) {
        plan_free(pl);
        return GSX_E_LAUNCH;
    }
    *plan = pl;
    return GSX_OK;
}

int gsx_batch_plan_launch(gsx_batch_plan* plan, void* stream) {
    if (!plan) return GSX_E_INVALID;
    return gsx::batch_enqueue(plan, gsx::as_stream(stream));
}

int gsx_batch_plan_launch_graph(gsx_batch_plan* plan, void* stream) {
    if (!plan) return GSX_E_INVALID;
    if (!plan->exec) {
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return GSX_E_LAUNCH;
        int st = GSX_OK;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) st = GSX_E_LAUNCH;
        if (st == GSX_OK) st = gsx::batch_enqueue(plan, cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (st == GSX_OK && e != hipSuccess) st = GSX_E_LAUNCH;
        if (st == GSX_OK && hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess)
            st = GSX_E_LAUNCH;
        (void)hipStreamDestroy(cs);
        if (st != GSX_OK) {
            if (g) (void)hipGraphDestroy(g);
            plan->exec = nullptr;
            return st;
        }
        plan->graph = g;
    }
    return hipGraphLaunch(plan->exec, gsx::as_stream(stream)) == hipSuccess ? GSX_OK : GSX_E_LAUNCH;
}

int64_t gsx_batch_plan_tiles(const gsx_batch_plan* plan) { return plan ? plan->ntiles : 0; }

void gsx_batch_plan_destroy(gsx_batch_plan* plan) { gsx::plan_free(plan); }

}  // extern "C"

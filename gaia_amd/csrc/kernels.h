// Host-callable launchers of the HIP kernels (all asynchronous on `stream`).
#pragma once
#include <hip/hip_runtime.h>

#include "device_types.h"

namespace gvbd {

// Lanes per vertex (1, 2, 4, 8, 16) and block size (64, 128, 256) of the group solve.
bool valid_schedule(int lpv, int block);
// whether an LDS-staged solve variant exists for (lpv, block, records per chunk); 0 is always valid
bool valid_stage(int lpv, int block, int stage);

// One parallel group: item i covers `count` consecutive colour-major slots of one scene,
// tet or cloth by item kind. Tet friction applies when iteration >= frictionStartIter
// (VBDPhysics.cpp:1610). `contact` selects the variant with contact forces and bound truncation
// (scenes with collision handling; stage must then be 0). `xcd` != 0 maps each XCD's blocks to one
// contiguous range of the items (results unchanged).
hipError_t launch_group_solve(const SceneView* scenes, const WorkItem* items, int nitems, int iteration, int lpv,
                              int block, int stage, bool contact, int xcd, hipStream_t stream);
// Device arguments of the persistent work-queue solve of a plan.
struct QueueArgs {
    const SceneView* scenes;
    const QueueEntry* queue;
    int nq;
    int* tickets;   // one ticket counter per launch (epoch) between counter resets
    int* progress;  // one completion counter per scene
    int* status;    // bit 0: a dependency wait exceeded spinLimit
    int spinLimit;
};
// One iteration of every scene of the plan (queue_solve_kernel) on `grid` persistent blocks;
// `epoch` counts the queue launches since the counters were last reset.
hipError_t launch_queue_solve(const QueueArgs& a, int iteration, int epoch, int grid, int lpv, int block,
                              hipStream_t stream);
// Resident blocks per CU of the queue kernel for a schedule (grid sizing).
int queue_blocks_per_cu(int lpv, int block);
// counters[0..n) = 0.
hipError_t launch_counter_reset(int* counters, int n, hipStream_t stream);
// Copies pnext -> pos for the cloth slots of `items` (after the group's solves).
hipError_t launch_cloth_commit(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream);
// Items cover vertex ranges [slotBegin, slotBegin + count), count <= 256.
hipError_t launch_initial_step(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream);
// deformed vertices of one phase (items of at most 256 DeformVert each): cloth meshes' before the
// initial step, tet meshes' after it
hipError_t launch_deform(const SceneView* scenes, const WorkItem* items, int nitems, bool cloth, hipStream_t stream);
hipError_t launch_update_velocity(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream);
// Chebyshev acceleration over the per-vertex items: history init after the initial step, and the
// accelerated update after iteration `iteration` (scenes without use_accelerator are skipped).
hipError_t launch_accel_init(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream);
hipError_t launch_accel_apply(const SceneView* scenes, const WorkItem* items, int nitems, int iteration,
                              hipStream_t stream);

// Convergence evaluation of one scene (VBDPhysics::evaluateConvergenceGPU): per-block partial
// sums of |force| over the tet slots, of m |y - x|^2 over the tet vertices `tetVerts` and of the
// elastic energy over `tets`; convergence_blocks(n) blocks of partials each.
int convergence_blocks(int n);
hipError_t launch_convergence(const SceneView& view, int nslots, const TetEnergyRec* tets, int nt,
                              const int* tetVerts, int ntv, double* forcePartial, double* inertiaPartial,
                              double* energyPartial, hipStream_t stream);

// Allocation-free evaluation: partials and totals live in a per-GPU file-scope device array of
// kConvergenceScratchDoubles doubles (1 MiB). convergence_scratch_init(device) looks up its address
// on `device`, which must be current, once per device; use only that address on the host side.
// The last 3 doubles receive the totals of launch_convergence_final_sum.
constexpr int kConvergenceScratchDoubles = 131072;
constexpr int kMaxDevices = 64;
hipError_t convergence_scratch_init(int device, double** scratch);
// out[k] = sum of partial[off_k .. off_k + n_k) in ascending order, k = 0, 1, 2
hipError_t launch_convergence_final_sum(const double* partial, int off0, int n0, int off1, int n1, int off2, int n2,
                                        double* out, hipStream_t stream);

// dst[i] = src[i] for n float4 elements (bandwidth / profiler calibration).
hipError_t launch_stream_copy(const float4* src, float4* dst, int64_t n, hipStream_t stream);

}  // namespace gvbd

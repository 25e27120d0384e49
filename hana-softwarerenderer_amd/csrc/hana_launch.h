// hana_launch.h -- host-callable launch wrappers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hana_types.h"

namespace hana {

hipError_t launch_vertex(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, hipStream_t s);
hipError_t launch_setup(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, bool small_tris,
                        hipStream_t s);
hipError_t launch_clip(const PassDesc* d_passes, const Workspace& ws, hipStream_t s);
int scan_chunks(int ntiles);
hipError_t launch_scan(const Workspace& ws, hipStream_t s);
// small_tris: every pass averages at most 4 pixels per face (smaller tile table)
hipError_t launch_bin(const PassDesc* d_passes, const Workspace& ws, int nseg, bool small_tris, hipStream_t s);
// one workgroup per tile of dependency group `group` (ws.tile_shift selects the tile size)
hipError_t launch_raster(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, int group,
                         hipStream_t s);
// the persistent raster over every pass of the sequence: `workgroups` workgroups take tiles from a queue
hipError_t launch_raster_queue(const PassDesc* d_passes, int npasses, const Workspace& ws, int workgroups,
                               hipStream_t s);
hipError_t launch_fill32(uint32_t* dst, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_present(const uint32_t* src, uint32_t* dst, int w, int h, hipStream_t s);

}  // namespace hana

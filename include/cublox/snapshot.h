/* Map snapshot format, version 1 (no reference counterpart: the reference
 * keeps no persistent map state). Little-endian; a 120-byte header, then
 * payload_bytes of payload:
 *   encoding 0 (dense):  float occupancy[voxel_num] in hash order;
 *   encoding 1 (sparse): uint32 count, uint32 0, int32 hash[count] ascending,
 *                        float value[count] — the voxels whose bit pattern
 *                        is nonzero (every other voxel is 0.0f).
 * The header carries everything OccupancyGrid needs to continue from the
 * snapshot: geometry (checked on load), window origin, first-run flag,
 * recenter threshold, log-odds parameters and the ray-cast range. */
#ifndef CUBLOX_SNAPSHOT_H
#define CUBLOX_SNAPSHOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUBLOX_SNAPSHOT_VERSION 1u
#define CUBLOX_SNAPSHOT_DENSE 0u
#define CUBLOX_SNAPSHOT_SPARSE 1u

typedef struct {
  char magic[8]; /* "CBLXSNAP" */
  uint32_t version;
  uint32_t encoding;
  int32_t half_map_size[3];
  float resolution;
  int32_t origin_at_center;
  int32_t has_threshold;
  float recenter_threshold;
  int32_t origin_i[3];
  float origin_f[3];
  int32_t first_run;
  float l_hit, l_miss, l_min, l_max, l_free, l_occupied, l_unknown;
  float max_raycast_range;
  int32_t voxel_num;
  uint32_t reserved;
  uint64_t payload_bytes;
} cublox_snapshot_header;

#ifdef __cplusplus
static_assert(sizeof(cublox_snapshot_header) == 120, "snapshot header is 120 bytes");
}
#endif

#endif /* CUBLOX_SNAPSHOT_H */

/*
 * noalloc.h -- the static device scratch behind arcx_step_device_noalloc
 * (noalloc.hip).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/arc_gpu.h"

/* Bytes of the static scratch per GPU (22 MiB): the largest contact lists of
 * any register-resident form of the quadruped for ARCX_NOALLOC_MAX_WORLDS
 * worlds (the four-lane forms' 4 x 29 slots of 6 floats = 2,784 bytes per
 * world; two lanes per world need 98 slots, one lane 89), the shared-root
 * form's lists and anchors for a block's 256 idle lanes (29 x 6 + 126 floats
 * per lane = 300 KiB), and a 256-byte work counter.  The entry point computes
 * the exact need of each sim from its model and refuses one that exceeds
 * this. */
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)ARCX_NOALLOC_MAX_WORLDS * 2784 + (size_t)300 * 1024 + 256;

/* Looks up the scratch's device address on `device` (which must be the
 * current device) once and keeps it; later calls return at once. */
hipError_t noalloc_scratch_init(int device);
/* the address noalloc_scratch_init kept for `device`, or null */
void *noalloc_scratch(int device);

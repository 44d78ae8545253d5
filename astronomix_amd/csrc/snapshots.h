// Snapshot rows and the snapshot schedule shared by the finite-volume and
// finite-difference step loops (time_integration.py:404-610, 671-678, 736-745).
#pragma once

#include <math.h>

#include "common.h"

namespace amx {
namespace snap {

// One row of SnapshotData at index: time point, state copy and diagnostics.
amx_status record(const amx_diag_desc& d, const float* state, float t, const amx_snapshot_out& o, int index,
                  void* work, hipStream_t st);

size_t work_bytes();

// Checks the spec and the output pointers against the layout d.
amx_status validate(const amx_snapshot_spec* spec, const amx_snapshot_out* out, const amx_diag_desc& d);

// time >= current_checkpoint * t_end / num_snapshots, in float32 as the
// reference evaluates it (int32 * float32, then / int) (time_integration.py:559-566)
inline bool due_uniform(int cc, float t, float t_end, int n) {
  const float thr = ((float)cc * t_end) / (float)n;
  return t >= thr;
}

// snapshot_timepoints[current_checkpoint] with the index clamped into range,
// as a JAX gather clamps out-of-bounds indices (time_integration.py:545, 673)
inline float timepoint(const amx_snapshot_spec* s, int cc) {
  const int i = cc < s->num_timepoints ? cc : s->num_timepoints - 1;
  return s->timepoints[i];
}

// |time - timepoint| < 1e-12 in float32 (time_integration.py:541-548, 738-740)
inline bool near(float a, float b) {
  const float d = a - b;
  return fabsf(d) < 1e-12f;
}

// The per-run snapshot driver of a step loop: decides at the start of each
// step (and after the last one with timepoints) whether a row is due, and
// records it or calls the user callback.
struct Driver {
  const amx_snapshot_spec* spec;
  amx_snapshot_out* out;
  amx_diag_desc desc;
  void* work;
  float t_end;
  int cc = 0;

  bool active() const { return spec != nullptr; }
  bool recording() const { return spec && !spec->callback; }
  // the timepoint dt clamp applies in recording mode only (time_integration.py:671-674)
  bool clamp() const { return recording() && spec->use_timepoints; }
  float clamp_target() const { return timepoint(spec, cc); }

  amx_status take(float t, const float* state, hipStream_t st) {
    if (recording()) {
      if (cc < spec->num_snapshots) {
        const amx_status s = record(desc, state, t, *out, cc, work, st);
        if (s != AMX_OK) return s;
      }
    } else {
      AMX_HIPCHK(hipStreamSynchronize(st));
      if (spec->callback(spec->user, t, cc, state) != 0)
        return amx_fail(AMX_ERR_INVALID, "the snapshot callback stopped the run%s");
    }
    ++cc;
    return AMX_OK;
  }

  // at the start of a step, with t the time before it
  amx_status begin(float t, const float* state, hipStream_t st) {
    if (!spec) return AMX_OK;
    const bool due = clamp() ? near(t, timepoint(spec, cc)) : due_uniform(cc, t, t_end, spec->num_snapshots);
    return due ? take(t, state, st) : AMX_OK;
  }

  // after a step, with t the time after it (timepoint mode records t_end)
  amx_status end(float t, const float* state, hipStream_t st) {
    if (!clamp()) return AMX_OK;
    return near(t, t_end) ? take(t, state, st) : AMX_OK;
  }

  void finish(long steps) {
    if (!out) return;
    out->current_checkpoint = cc;
    out->num_iterations = steps;
  }
};

}  // namespace snap
}  // namespace amx

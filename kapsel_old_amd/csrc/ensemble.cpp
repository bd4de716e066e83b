// Ensembles of independent SPM boxes on one GPU (include/spm_ensemble.h).
//
// Each box is an ordinary context.  One step of every box is captured into
// a HIP graph whose branches ("lanes") each run a share of the boxes one
// after another; the shares come from a longest-processing-time-first
// split of the boxes' step times.  Replaying the graph advances the whole
// ensemble by a step with one launch.  The thermostat reads normals drawn
// ahead into a per-box device pool (spm::pool_fill), so a replay needs no
// host work.  Nothing in a captured step synchronises with the host or
// records a memset node (counters and fields are cleared by kernels).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "spm_internal.h"
#include "../../include/spm_ensemble.h"

struct spm_ensemble {
  int device = 0;
  std::vector<spm_ctx *> box;
  std::vector<hipStream_t> own_stream;  // each box's own stream, restored at destroy
  std::vector<double> cost;             // estimated, then measured, ms per step
  std::vector<double> box_ms;           // measured (0 until autotuned)
  hipStream_t master = nullptr;
  hipStream_t lanes[SPM_ENSEMBLE_MAX_LANES] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[SPM_ENSEMBLE_MAX_LANES] = {};
  hipEvent_t copy_done[2] = {};  // the pool copies of each half of the ring
  int nlanes = 1;
  std::vector<std::vector<int>> split;  // boxes of each lane
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int exec_lanes = 0;  // lane count the instantiated graph was built for
  double tune_ms[4] = {0, 0, 0, 0};
  std::string err;
};

namespace {

int fail(spm_ensemble *e, int rc, const std::string &msg) {
  e->err = msg;
  spm::global_error() = msg;
  return rc;
}

#define EH(call)                                                         \
  do {                                                                   \
    hipError_t h_ = (call);                                              \
    if (h_ != hipSuccess)                                                \
      return fail(e, SPM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(h_)); \
  } while (0)

// a box's error becomes the ensemble's, prefixed with the box index
int box_fail(spm_ensemble *e, int b, int rc) {
  char pre[32];
  std::snprintf(pre, sizeof pre, "box %d: ", b);
  return fail(e, rc, pre + e->box[b]->err);
}

// step-cost estimate before anything is measured: transforms ~ N log N
// over about 20 field transforms, plus the particle stencil work
double estimate_cost(const spm_ctx *c) {
  const double n = (double)c->grid.nx * c->grid.ny * c->grid.nz;
  return 20.0 * n * std::log2(n) * 1e-9 + 40.0 * (double)c->np * c->n_stencil * 1e-9 + 0.05;
}

// longest processing time first onto `lanes` bins
std::vector<std::vector<int>> lpt_split(const std::vector<double> &cost, int lanes) {
  std::vector<int> order(cost.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<std::vector<int>> bins(lanes);
  std::vector<double> load(lanes, 0.0);
  for (int b : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bins[l].push_back(b);
    load[l] += cost[b];
  }
  for (auto &bin : bins) std::sort(bin.begin(), bin.end());
  return bins;
}

void drop_graph(spm_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->exec_lanes = 0;
}

int set_lanes(spm_ensemble *e, int lanes) {
  lanes = std::max(1, std::min(lanes, std::min((int)e->box.size(), SPM_ENSEMBLE_MAX_LANES)));
  if (lanes != e->nlanes || e->split.empty()) {
    e->nlanes = lanes;
    e->split = lpt_split(e->cost, lanes);
    drop_graph(e);
  }
  return 0;
}

int wait_all(spm_ensemble *e) {
  for (int l = 0; l < SPM_ENSEMBLE_MAX_LANES; ++l) EH(hipStreamSynchronize(e->lanes[l]));
  EH(hipStreamSynchronize(e->master));
  return 0;
}

int check_boxes(spm_ensemble *e) {
  for (size_t b = 0; b < e->box.size(); ++b) {
    const spm_ctx *c = e->box[b];
    if (c->electro || c->le || c->timing) {
      char m[128];
      std::snprintf(m, sizeof m, "box %zu: ensembles run the plain Navier-Stokes step (no electrolyte, "
                    "shear or timing)", b);
      return fail(e, SPM_ERR_UNSUPPORTED, m);
    }
    // a network's input and output buffers serve one box at a time, and the
    // boxes of different lanes run at once
    for (size_t o = 0; o < b && c->closure; ++o)
      if (e->box[o]->closure == c->closure) {
        char m[128];
        std::snprintf(m, sizeof m, "boxes %zu and %zu share one closure network", o, b);
        return fail(e, SPM_ERR_UNSUPPORTED, m);
      }
  }
  return 0;
}

// one step of every box, launched from the host on the lanes; with
// box_events the boxes run one after another on lane 0 with an event pair
// around each (their step times)
int eager_step(spm_ensemble *e, bool pool, std::vector<hipEvent_t> *box_events) {
  const int nb = (int)e->box.size();
  for (int l = 0; l < e->nlanes; ++l) {
    for (int b : e->split[l]) {
      spm_ctx *c = e->box[b];
      c->stream = box_events ? e->lanes[0] : e->lanes[l];
      c->pool_mode = pool && spm::box_uses_noise(c);
      if (box_events) EH(hipEventRecord((*box_events)[2 * b], c->stream));
      if (int rc = spm::box_step_phase(c, 1)) {
        c->pool_mode = false;
        return box_fail(e, b, rc);
      }
      if (box_events) EH(hipEventRecord((*box_events)[2 * b + 1], c->stream));
      c->pool_mode = false;
    }
  }
  for (int b = 0; b < nb; ++b) {
    e->box[b]->time += e->box[b]->prm.dt;
    e->box[b]->ts++;
  }
  return 0;
}

int capture(spm_ensemble *e) {
  drop_graph(e);
  EH(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  int rc = 0, bad = -1;
  if (hipEventRecord(e->fork, e->master) != hipSuccess) rc = SPM_ERR_HIP;
  for (int l = 0; l < e->nlanes && !rc; ++l) {
    if (hipStreamWaitEvent(e->lanes[l], e->fork, 0) != hipSuccess) {
      rc = SPM_ERR_HIP;
      break;
    }
    for (int b : e->split[l]) {
      spm_ctx *c = e->box[b];
      c->stream = e->lanes[l];
      c->pool_mode = spm::box_uses_noise(c);
      rc = spm::box_step_phase(c, 1);
      c->pool_mode = false;
      if (rc) {
        bad = b;
        break;
      }
    }
    if (!rc && (hipEventRecord(e->join[l], e->lanes[l]) != hipSuccess ||
                hipStreamWaitEvent(e->master, e->join[l], 0) != hipSuccess))
      rc = SPM_ERR_HIP;
  }
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(e->master, &g);
  if (rc || end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (bad >= 0) return box_fail(e, bad, rc);
    return fail(e, SPM_ERR_HIP, std::string("graph capture of the ensemble step failed: ") +
                                    hipGetErrorString(end != hipSuccess ? end : hipErrorUnknown));
  }
  e->graph = g;
  EH(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  e->exec_lanes = e->nlanes;
  return 0;
}

// nsteps steps of every box with ts > 0, replayed (or host-launched) on the
// master stream, optionally timed (ms: device time between events around
// the steps).  The noise pools are rings of 2 SPM_ENSEMBLE_CHUNK slots: the
// host draws chunk i + 1 into one half while the device reads chunk i from
// the other; a half's staging is rewritten only after the copy that last
// read it has finished (copy_done), and its device slots only by a copy
// queued behind the replays that read them.
int run_steps(spm_ensemble *e, int nsteps, float *ms) {
  const int nb = (int)e->box.size();
  if (int rc = wait_all(e)) return rc;
  for (int b = 0; b < nb; ++b) {
    spm_ctx *c = e->box[b];
    c->stream = e->master;
    if (int rc = spm::pool_reset(c)) return box_fail(e, b, rc);
  }
  if (e->use_graph && (!e->exec || e->exec_lanes != e->nlanes))
    if (int rc = capture(e)) return rc;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (ms) {
    EH(hipEventCreate(&t0));
    EH(hipEventCreate(&t1));
  }
  bool timing_started = false;
  for (int done = 0, chunk = 0; done < nsteps; ++chunk) {
    const int k = std::min(nsteps - done, SPM_ENSEMBLE_CHUNK);
    const int half = chunk & 1;
    EH(hipEventSynchronize(e->copy_done[half]));
    for (int b = 0; b < nb; ++b) {
      spm_ctx *c = e->box[b];
      c->stream = e->master;
      if (int rc = spm::pool_fill(c, k, half * SPM_ENSEMBLE_CHUNK)) return box_fail(e, b, rc);
    }
    EH(hipEventRecord(e->copy_done[half], e->master));
    if (ms && !timing_started) {
      EH(hipEventRecord(t0, e->master));
      timing_started = true;
    }
    for (int s = 0; s < k; ++s) {
      if (e->use_graph) {
        EH(hipGraphLaunch(e->exec, e->master));
        for (int b = 0; b < nb; ++b) {
          e->box[b]->time += e->box[b]->prm.dt;
          e->box[b]->ts++;
        }
      } else {
        // lanes wait for the pool copies on the master, the master for the lanes
        EH(hipEventRecord(e->fork, e->master));
        for (int l = 0; l < e->nlanes; ++l) EH(hipStreamWaitEvent(e->lanes[l], e->fork, 0));
        if (int rc = eager_step(e, true, nullptr)) return rc;
        for (int l = 0; l < e->nlanes; ++l) {
          EH(hipEventRecord(e->join[l], e->lanes[l]));
          EH(hipStreamWaitEvent(e->master, e->join[l], 0));
        }
      }
    }
    done += k;
  }
  if (ms) {
    EH(hipEventRecord(t1, e->master));
    EH(hipEventSynchronize(t1));
    EH(hipEventElapsedTime(ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  return wait_all(e);
}

// one step of every box launched from the host, one box after another on
// lane 0, the thermostat drawing as spm_step does: the step a box at
// ts == 0 takes first (the Euler start of the AB2 integrators), and any box
// past it alongside so that all advance alike
int plain_step_all(spm_ensemble *e) {
  if (int rc = wait_all(e)) return rc;
  for (size_t b = 0; b < e->box.size(); ++b) {
    spm_ctx *c = e->box[b];
    c->stream = e->lanes[0];
    if (int rc = spm::box_step_phase(c, 1)) return box_fail(e, (int)b, rc);
    c->time += c->prm.dt;
    c->ts++;
  }
  return wait_all(e);
}

bool any_unstarted(const spm_ensemble *e) {
  for (const spm_ctx *c : e->box)
    if (c->ts == 0) return true;
  return false;
}

int phase_all(spm_ensemble *e, int phase) {
  for (size_t b = 0; b < e->box.size(); ++b) {
    spm_ctx *c = e->box[b];
    c->stream = e->lanes[0];
    if (int rc = spm::box_step_phase(c, phase)) return box_fail(e, (int)b, rc);
  }
  return wait_all(e);
}

// nsteps steps of every box, as spm_step(box, nsteps) would take them
int ensemble_step(spm_ensemble *e, int nsteps, float *ms = nullptr) {
  if (nsteps <= 0) return 0;
  if (int rc = check_boxes(e)) return rc;
  if (int rc = phase_all(e, 0)) return rc;
  int left = nsteps;
  if (any_unstarted(e)) {
    if (int rc = plain_step_all(e)) return rc;
    --left;
  }
  if (left > 0)
    if (int rc = run_steps(e, left, ms)) return rc;
  if (int rc = phase_all(e, 2)) return rc;
  for (size_t b = 0; b < e->box.size(); ++b) e->box[b]->stream = e->own_stream[b];
  return 0;
}

}  // namespace

extern "C" {

int spm_ensemble_create(int nbox, const spm_params *params, spm_ensemble **out) {
  spm::global_error().clear();
  if (!out || !params || nbox < 1) {
    spm::global_error() = "need at least one box and an output pointer";
    return SPM_ERR_INVALID;
  }
  *out = nullptr;
  for (int b = 1; b < nbox; ++b)
    if (params[b].device != params[0].device) {
      spm::global_error() = "every box of an ensemble must be on the same device";
      return SPM_ERR_INVALID;
    }
  spm_ensemble *e = new (std::nothrow) spm_ensemble();
  if (!e) return SPM_ERR_OOM;
  e->device = params[0].device;
  for (int b = 0; b < nbox; ++b) {
    spm_ctx *c = nullptr;
    const int rc = spm_create(&params[b], &c);
    if (rc) {
      char pre[32];
      std::snprintf(pre, sizeof pre, "box %d: ", b);
      const std::string msg = pre + spm::global_error();
      spm_ensemble_destroy(e);
      spm::global_error() = msg;
      return rc;
    }
    e->box.push_back(c);
    e->own_stream.push_back(c->stream);
  }
  int rc = 0;
  if (hipSetDevice(e->device) != hipSuccess || hipStreamCreate(&e->master) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->copy_done[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->copy_done[1], hipEventDisableTiming) != hipSuccess)
    rc = SPM_ERR_HIP;
  for (int h = 0; h < 2 && !rc; ++h)  // both halves of the pools start free
    if (hipEventRecord(e->copy_done[h], e->master) != hipSuccess) rc = SPM_ERR_HIP;
  for (int l = 0; l < SPM_ENSEMBLE_MAX_LANES && !rc; ++l)
    if (hipStreamCreate(&e->lanes[l]) != hipSuccess ||
        hipEventCreateWithFlags(&e->join[l], hipEventDisableTiming) != hipSuccess)
      rc = SPM_ERR_HIP;
  for (int b = 0; b < nbox && !rc; ++b) {
    spm_ctx *c = e->box[b];
    e->cost.push_back(estimate_cost(c));
    rc = spm::pool_setup(c, 2 * SPM_ENSEMBLE_CHUNK);
    if (!rc) rc = spm::tune_mixed_transforms(c);
    if (rc) {
      const std::string msg = "box " + std::to_string(b) + ": " + c->err;
      spm_ensemble_destroy(e);
      spm::global_error() = msg;
      return rc;
    }
  }
  if (rc) {
    spm_ensemble_destroy(e);
    spm::global_error() = "creating the ensemble's streams failed";
    return rc;
  }
  e->box_ms.assign(nbox, 0.0);
  set_lanes(e, 4);
  *out = e;
  return SPM_OK;
}

int spm_ensemble_destroy(spm_ensemble *e) {
  if (!e) return SPM_OK;
  (void)hipSetDevice(e->device);
  for (int l = 0; l < SPM_ENSEMBLE_MAX_LANES; ++l)
    if (e->lanes[l]) (void)hipStreamSynchronize(e->lanes[l]);
  if (e->master) (void)hipStreamSynchronize(e->master);
  drop_graph(e);
  for (size_t b = 0; b < e->box.size(); ++b) {
    e->box[b]->stream = e->own_stream[b];
    spm_destroy(e->box[b]);
  }
  for (int l = 0; l < SPM_ENSEMBLE_MAX_LANES; ++l) {
    if (e->lanes[l]) (void)hipStreamDestroy(e->lanes[l]);
    if (e->join[l]) (void)hipEventDestroy(e->join[l]);
  }
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (hipEvent_t ev : e->copy_done)
    if (ev) (void)hipEventDestroy(ev);
  if (e->master) (void)hipStreamDestroy(e->master);
  delete e;
  return SPM_OK;
}

spm_ctx *spm_ensemble_box(spm_ensemble *e, int b) {
  if (!e || b < 0 || b >= (int)e->box.size()) return nullptr;
  return e->box[b];
}

int spm_ensemble_size(const spm_ensemble *e) { return e ? (int)e->box.size() : -1; }

int spm_ensemble_step(spm_ensemble *e, int nsteps) {
  if (!e || nsteps < 0) return SPM_ERR_INVALID;
  e->err.clear();
  if (hipSetDevice(e->device) != hipSuccess) return fail(e, SPM_ERR_HIP, "hipSetDevice failed");
  const int rc = ensemble_step(e, nsteps);
  if (rc) {
    (void)wait_all(e);
    for (size_t b = 0; b < e->box.size(); ++b) e->box[b]->stream = e->own_stream[b];
  }
  return rc;
}

int spm_ensemble_set_graph(spm_ensemble *e, int graph) {
  if (!e) return SPM_ERR_INVALID;
  e->use_graph = graph != 0;
  return SPM_OK;
}

int spm_ensemble_autotune(spm_ensemble *e, int steps_per_candidate) {
  if (!e || steps_per_candidate < 1 || steps_per_candidate > SPM_ENSEMBLE_CHUNK) return SPM_ERR_INVALID;
  e->err.clear();
  if (hipSetDevice(e->device) != hipSuccess) return fail(e, SPM_ERR_HIP, "hipSetDevice failed");
  if (int rc = check_boxes(e)) return rc;
  const int nb = (int)e->box.size();
  auto restore = [&](int rc) {
    (void)wait_all(e);
    for (int b = 0; b < nb; ++b) e->box[b]->stream = e->own_stream[b];
    return rc;
  };
  if (int rc = phase_all(e, 0)) return restore(rc);
  // the Euler start, where needed, then one step of every box alone on one
  // stream with events around each: the measured costs of the split
  if (any_unstarted(e))
    if (int rc = plain_step_all(e)) return restore(rc);
  {
    for (int b = 0; b < nb; ++b) {  // one step of noise in slot 0
      spm_ctx *c = e->box[b];
      c->stream = e->master;
      int rc = spm::pool_reset(c);
      if (!rc) rc = spm::pool_fill(c, 1, 0);
      if (rc) return restore(box_fail(e, b, rc));
    }
    if (int rc = wait_all(e)) return restore(rc);
    std::vector<hipEvent_t> ev(2 * nb, nullptr);
    for (auto &x : ev) EH(hipEventCreate(&x));
    int rc = eager_step(e, true, &ev);
    if (!rc) rc = wait_all(e);
    for (int b = 0; b < nb && !rc; ++b) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[2 * b], ev[2 * b + 1]) != hipSuccess) rc = SPM_ERR_HIP;
      e->box_ms[b] = ms;
      e->cost[b] = ms;
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
    if (rc) return restore(rc == SPM_ERR_HIP ? fail(e, rc, "timing the boxes failed") : rc);
  }
  static const int cand[4] = {1, 2, 4, 8};
  int best = -1;
  for (int k = 0; k < 4; ++k) {
    e->tune_ms[k] = 0.0;
    if (cand[k] > nb) continue;
    e->split.clear();
    set_lanes(e, cand[k]);
    float ms = 0.f;
    if (int rc = run_steps(e, 1, nullptr)) return restore(rc);  // capture + first replay, untimed
    if (int rc = run_steps(e, steps_per_candidate, &ms)) return restore(rc);
    e->tune_ms[k] = ms / steps_per_candidate;
    if (best < 0 || e->tune_ms[k] < e->tune_ms[best]) best = k;
  }
  e->split.clear();
  set_lanes(e, cand[best]);
  if (int rc = phase_all(e, 2)) return restore(rc);
  return restore(0);
}

int spm_ensemble_tuning(const spm_ensemble *e, int *lanes, double ms[4], double *box_ms) {
  if (!e || !lanes || !ms) return SPM_ERR_INVALID;
  *lanes = e->nlanes;
  for (int k = 0; k < 4; ++k) ms[k] = e->tune_ms[k];
  if (box_ms)
    for (size_t b = 0; b < e->box.size(); ++b) box_ms[b] = e->box_ms[b];
  return SPM_OK;
}

int spm_autotune_report(char *buf, int len) {
  const std::string r = spm::tune_report();
  if (buf && len > 0) {
    const size_t n = std::min((size_t)len - 1, r.size());
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return (int)r.size();
}

const char *spm_ensemble_last_error(const spm_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

}  // extern "C"

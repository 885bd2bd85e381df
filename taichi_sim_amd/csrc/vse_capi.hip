// vse_capi.hip -- C-ABI of libvse (declared in include/vse.h).
//
// Owns the device state of one simulation, turns the reference's host tables
// into the kernels' compact forms (CPML bands, source list, receiver CSR) and
// runs the step loop as hipGraphs of up to `check_every` steps on a private
// stream, checking the running |v|^2 max between graphs.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#define VSE_HOST_SIDE
#include "vse_internal.h"
#include "../../include/vse.h"

namespace {

// axis of the derivative each psi slot belongs to
const int kSlotAxis[P_COUNT] = {0, 1, 2, 0, 1, 0, 2, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2};

}  // namespace

struct vse_sim {
  vse_desc d{};
  int device = 0;
  hipStream_t stream = nullptr;
  long ncell = 0;
  std::string err;
  std::string devname;
  size_t bytes = 0;
  // device allocations as (address handed out, address hipMalloc returned)
  std::vector<std::pair<void *, void *>> allocs;
  static std::atomic<unsigned> g_allocs;    // device arrays allocated so far (the base skew)
  // mechanism stride of the memory variables in floats: mechanism l of
  // component c is r[c] + l * rs (ncell plus a skew, see alloc())
  long rs = 0;

  float *field[9] = {};
  float *rho = nullptr, *cp = nullptr, *cs = nullptr;
  bool have_material = false;
  float *prof[3][8] = {};     // a, b, k, a_h, b_h, k_h, 1/k, 1/k_h
  int math = VSE_MATH_STRICT;
  AxisProf ax[3]{};
  bool have_cpml[3] = {false, false, false};
  float *psi[P_COUNT] = {};
  bool psi_ready = false;
  float *r[6] = {};
  ZenerK zk{};
  unsigned *ctrl = nullptr;       // [0] = v2max bits, [1] = step counter
  unsigned *host_ctrl = nullptr;  // pinned mirror
  // 2D: running max of |v|^2 per velocity workgroup, merged into ctrl[0]
  // (fold_v2) before the host reads it
  unsigned *v2slots = nullptr;
  int n_slots = 0;
  // sources
  int n_src_pts = 0, src_cols = 1;
  int *src_cell = nullptr, *src_col = nullptr, *src_map = nullptr;
  int *src_rowflag = nullptr;     // 2D: per x-row, 1 if it holds a source cell
  float *src = nullptr;
  // receivers
  int n_rec = 0, rec_gate = 0;
  int rec_channel = VSE_F_SYY;    // field the receivers sample (VSE_F_*)
  int *rec_beg = nullptr, *rec_cell = nullptr, *rec_delay = nullptr;
  float *traces = nullptr;
  // 3D sources by (x, y) row (the fused step's table)
  int *src_rbeg = nullptr, *src_rz = nullptr, *src_rcol = nullptr;
  // launch shapes of the 3D stress / velocity kernels, and of the fused step
  KTune tune[3] = {{4, 0, 0, 0}, {4, 0, 8, 1}, {8, 0, 32, 2}};
  // schedule of the 3D step; the fused one needs a second set of sigma, v and
  // velocity-pass psi buffers (set 1; set 0 is field[] / psi[])
  int sched = VSE_SCHED_TWO_PASS;
  float *alt[9] = {};
  float *psi_alt[P_COUNT] = {};
  bool alt_ready = false;
  // cached graphs (two-pass steps; fused steps)
  hipGraphExec_t gexec = nullptr;
  int gexec_steps = 0;
  hipGraphExec_t fgexec = nullptr;
  int fgexec_steps = 0;
  // handles built on this simulation; vse_destroy detaches them, so any
  // destroy order is safe (a detached handle reports VSE_ERR_STATE)
  std::vector<vse_batch *> batches;
  std::vector<vse_grad *> grads;

  int fail(int code, const std::string &m) { err = m; return code; }
  int hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return VSE_ERR_HIP;
  }
  // Zeroed device array of n elements.  Successive arrays start 256 B apart
  // modulo 4 KiB: the passes stream 10-30 arrays at the same cell index, and
  // with every base on the same 4 KiB alignment those streams meet in the
  // same HBM channels (measured on a 30-array in-place mix at 512^3: 5.0 TB/s
  // with aligned bases, 5.2 TB/s skewed).  The count runs over every
  // simulation of the process, so the same field of batch members that run
  // together does not share an alignment either.
  template <typename T>
  int alloc(T **p, size_t n) {
    constexpr size_t kSkewStep = 256, kSkewSpan = 16;
    const size_t nb = std::max<size_t>(n, 1) * sizeof(T);
    const size_t skew = kSkewStep * (g_allocs.fetch_add(1, std::memory_order_relaxed) % kSkewSpan);
    void *q = nullptr;
    int rc = hip(hipMalloc(&q, nb + kSkewStep * kSkewSpan), "hipMalloc");
    if (rc) return rc;
    void *u = (char *)q + skew;
    allocs.emplace_back(u, q);
    bytes += nb;
    *p = (T *)u;
    return hip(hipMemsetAsync(u, 0, nb, stream), "hipMemsetAsync");
  }
  void release(void *p) {
    if (!p) return;
    auto it = std::find_if(allocs.begin(), allocs.end(), [p](const std::pair<void *, void *> &a) { return a.first == p; });
    if (it == allocs.end()) return;
    (void)hipFree(it->second);
    allocs.erase(it);
  }
  // the device step counter (ctrl[1])
  int *ctrl_step() { return (int *)(ctrl + 1); }
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gexec_steps = 0;
    if (fgexec) (void)hipGraphExecDestroy(fgexec);
    fgexec = nullptr;
    fgexec_steps = 0;
  }
  long axis_len(int a) const { return a == 0 ? d.nx : (a == 1 ? d.ny : d.nz); }
  long psi_elems(int slot) const {
    const int a = kSlotAxis[slot];
    if (d.ndim == 2) return a == 0 ? (long)ax[0].nb * d.ny : (long)d.nx * ax[1].nb;
    if (a == 0) return (long)ax[0].nb * d.ny * d.nz;
    if (a == 1) return (long)d.nx * ax[1].nb * d.nz;
    return (long)d.nx * d.ny * ax[2].nb;
  }
  bool slot_used(int slot) const {
    if (d.ndim == 3) return true;
    for (int u : kPsiSlots2D)
      if (u == slot) return true;
    return false;
  }
};

std::atomic<unsigned> vse_sim::g_allocs{0};

static int check_handle(vse_sim *s) { return s ? 0 : VSE_ERR_ARG; }

// ---- scratch of vse_run_noalloc ----------------------------------------------
// Word 0: running max of |v|^2 (float bits), word 1: step counter, words
// 2 .. 2 + VSE_NOALLOC_SLOTS: the 2D velocity workgroups' |v|^2 slots.  Each
// GPU holds its own copy until the process ends.  In host code `scratch` is
// only a stand-in symbol: its device address, per device, is scratch_dev[]
// (hipGetSymbolAddress at the device's first vse_create).
#define VSE_SCRATCH_WORDS (2 + VSE_NOALLOC_SLOTS)
static __device__ __align__(256) unsigned scratch[VSE_SCRATCH_WORDS];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static unsigned *scratch_dev[kMaxDevices] = {};

// Looks up (once per device, with that device current) the device address of
// this GPU's scratch copy; 0 on success.
static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return VSE_ERR_ARG;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return 0;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VSE_ERR_HIP;
  scratch_dev[device] = (unsigned *)p;
  return 0;
}

static unsigned *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_dev[device];
}

// scr[0..1] = ctrl[0..1] (the handle's running max and step), scr[2 .. 2+n) = 0
__global__ __launch_bounds__(256) void k_scratch_begin(unsigned *scr, const unsigned *ctrl, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) scr[i] = ctrl[i];
  else if (i < 2 + n) scr[i] = 0u;
}

// ctrl[0..1] = scr[0..1]
__global__ __launch_bounds__(64) void k_scratch_end(const unsigned *scr, unsigned *ctrl) {
  if (threadIdx.x < 2) ctrl[threadIdx.x] = scr[threadIdx.x];
}

// vse_destroy(s) tells the handles built on s that it is gone (defined below
// with their types): a batch forgets the member, a gradient handle releases
// its device memory and reports VSE_ERR_STATE from then on.
static void batch_forget_member(vse_batch *b, vse_sim *s);
static void grad_detach(vse_grad *g);

// Waits for the stream when it goes out of scope.  A setter that queues
// host-to-device copies from its own local buffers declares one first, so that
// every return -- an error return included -- leaves no copy reading freed memory.
class SyncOnExit {
 public:
  explicit SyncOnExit(hipStream_t st) : st_(st) {}
  ~SyncOnExit() {
    if (st_) (void)hipStreamSynchronize(st_);
  }
  SyncOnExit(const SyncOnExit &) = delete;
  SyncOnExit &operator=(const SyncOnExit &) = delete;

 private:
  hipStream_t st_;
};

// Field ids a simulation holds (2D: vx, vy, sxx, syy, sxy; 3D: all nine).
static bool field_present(const vse_desc &d, int f) {
  if (f < 0 || f > 8) return false;
  return d.ndim == 3 || f == VSE_F_VX || f == VSE_F_VY || f == VSE_F_SXX || f == VSE_F_SYY || f == VSE_F_SXY;
}

static bool velocity_field(int f) { return f == VSE_F_VX || f == VSE_F_VY || f == VSE_F_VZ; }

// Makes a handle's device current for one C-ABI call and restores the
// caller's device on return, so a caller that switched devices between calls
// (torch.cuda.set_device, another library) still reaches the handle's GPU.
class OnDevice {
 public:
  explicit OnDevice(int dev) {
    if (hipGetDevice(&prev_) != hipSuccess || prev_ == dev || hipSetDevice(dev) != hipSuccess) prev_ = -1;
  }
  ~OnDevice() {
    if (prev_ >= 0) (void)hipSetDevice(prev_);
  }
  OnDevice(const OnDevice &) = delete;
  OnDevice &operator=(const OnDevice &) = delete;

 private:
  int prev_ = -1;
};

extern "C" {

int vse_abi_version(void) { return VSE_ABI_VERSION; }

int vse_create(const vse_desc *desc, vse_sim **out) {
  if (!desc || !out) return VSE_ERR_ARG;
  *out = nullptr;
  const vse_desc &d = *desc;
  if (d.ndim != 2 && d.ndim != 3) return VSE_ERR_ARG;
  if (d.ord < 2 || d.ord > VSE_MAX_ORD) return VSE_ERR_ARG;
  if (d.n_sls < 1 || d.n_sls > VSE_MAX_SLS) return VSE_ERR_ARG;
  if (d.variant != VSE_VARIANT_WGSL && d.variant != VSE_VARIANT_TAICHI) return VSE_ERR_ARG;
  if (d.ndim == 3 && d.variant != VSE_VARIANT_WGSL) return VSE_ERR_ARG;
  const int minn = 2 * d.ord + 1;
  if (d.nx < minn || d.ny < minn || (d.ndim == 3 && d.nz < minn)) return VSE_ERR_ARG;
  if (d.n_rows < 1 || !(d.dt > 0.f) || !(d.dx > 0.f) || !(d.dy > 0.f) || (d.ndim == 3 && !(d.dz > 0.f)))
    return VSE_ERR_ARG;
  // 32-bit buffer addressing: the 2D velocity kernels index whole fields
  // (byte offsets and ranges below 2^32), the 3D kernels index x-planes
  if (d.ndim == 2 && (long)d.nx * d.ny >= (1L << 30)) return VSE_ERR_ARG;
  if (d.ndim == 3 && (long)d.ny * d.nz >= (1L << 30)) return VSE_ERR_ARG;

  vse_sim *s = new vse_sim();
  s->d = d;
  if (d.ndim == 2) {
    s->d.nz = 1;
    s->tune[0] = s->tune[1] = KTune{4, 0, 0, 1};   // XCD-contiguous tile order
    s->tune[2] = KTune{4, 0, 16, 1};                // fused step: 16 x-rows per workgroup
  } else {
    s->rec_channel = VSE_F_SZZ;
  }
  s->ncell = (long)s->d.nx * s->d.ny * s->d.nz;
  int rc = s->hip(hipGetDevice(&s->device), "hipGetDevice");
  if (!rc && scratch_setup(s->device)) rc = s->fail(VSE_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
  if (!rc) rc = s->hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!rc) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) == hipSuccess)
      s->devname = (prop.name[0] ? std::string(prop.name) : std::string("AMD GPU")) + " (" + prop.gcnArchName +
                   ", " + std::to_string(prop.multiProcessorCount) + " CUs, " +
                   std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
  }
  for (int f = 0; f < 9 && !rc; f++) {
    const bool used = d.ndim == 3 || f == VSE_F_VX || f == VSE_F_VY || f == VSE_F_SXX || f == VSE_F_SYY ||
                      f == VSE_F_SXY;
    if (used) rc = s->alloc(&s->field[f], s->ncell);
  }
  // mechanisms 1.5 KiB apart modulo 4 KiB, for the channel spread alloc() gives
  // whole arrays
  s->rs = s->ncell + 384;
  if (!rc && d.visco && d.variant == VSE_VARIANT_WGSL) {
    for (int c = 0; c < 6 && !rc; c++)
      if (d.ndim == 3 || c == 0 || c == 1 || c == 3) rc = s->alloc(&s->r[c], (size_t)s->rs * d.n_sls);
  }
  if (!rc) rc = s->alloc(&s->ctrl, 2);
  if (!rc) rc = s->hip(hipHostMalloc((void **)&s->host_ctrl, 2 * sizeof(unsigned), hipHostMallocDefault),
                       "hipHostMalloc");
  if (!rc && d.ndim == 2) {
    // one slot per workgroup of the widest 2D velocity launch (one BX-row group per tile)
    const long n = vse_tiles2d(d.nx, d.ny, 1);
    if (n > 0x7fffffffL) rc = s->fail(VSE_ERR_ARG, "grid too large");
    else { s->n_slots = (int)n; rc = s->alloc(&s->v2slots, s->n_slots); }
  }
  // Zener constants per mechanism, float32 in the reference's operation order
  // (wgsl2d :881-894): phi = dt(1-alpha)(1/tau_sig)/sum_alpha, h = 0.5 dt/tau_sig,
  // m = 1/(1+h); identical to the per-cell values the reference recomputes.
  if (!rc) {
    const int L = d.n_sls;
    for (int l = 0; l < L; l++) {
      const float t_ep = d.tau[0 * L + l], t_sp = d.tau[1 * L + l];
      const float t_es = d.tau[2 * L + l], t_ss = d.tau[3 * L + l];
      const float inv_p = 1.0f / t_sp, inv_s = 1.0f / t_ss;
      const float al_p = t_ep / t_sp, al_s = t_es / t_ss;
      s->zk.phi_p[l] = d.dt * (1.0f - al_p) * inv_p / d.sum_alpha[0];
      s->zk.phi_s[l] = d.dt * (1.0f - al_s) * inv_s / d.sum_alpha[1];
      s->zk.h_p[l] = 0.5f * d.dt * inv_p;
      s->zk.h_s[l] = 0.5f * d.dt * inv_s;
      s->zk.m_p[l] = 1.0f / (1.0f + s->zk.h_p[l]);
      s->zk.m_s[l] = 1.0f / (1.0f + s->zk.h_s[l]);
    }
  }
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "create");
  if (rc) {
    vse_destroy(s);
    return rc;
  }
  *out = s;
  return VSE_OK;
}

int vse_set_material(vse_sim *s, const float *rho, const float *cp, const float *cs) {
  if (check_handle(s) || !rho || !cp || !cs) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  SyncOnExit sync(s->stream);
  int rc = 0;
  if (!s->rho) {
    rc = s->alloc(&s->rho, s->ncell);
    if (!rc) rc = s->alloc(&s->cp, s->ncell);
    if (!rc) rc = s->alloc(&s->cs, s->ncell);
    if (rc) return rc;
  }
  const size_t nb = s->ncell * sizeof(float);
  rc = s->hip(hipMemcpyAsync(s->rho, rho, nb, hipMemcpyHostToDevice, s->stream), "copy rho");
  if (!rc) rc = s->hip(hipMemcpyAsync(s->cp, cp, nb, hipMemcpyHostToDevice, s->stream), "copy cp");
  if (!rc) rc = s->hip(hipMemcpyAsync(s->cs, cs, nb, hipMemcpyHostToDevice, s->stream), "copy cs");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "set_material");
  if (!rc) s->have_material = true;
  return rc;
}

int vse_set_cpml(vse_sim *s, int axis, const float *a, const float *b, const float *k, const float *a_h,
                 const float *b_h, const float *k_h) {
  if (check_handle(s) || axis < 0 || axis >= s->d.ndim) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  const float *src[6] = {a, b, k, a_h, b_h, k_h};
  for (int i = 0; i < 6; i++)
    if (!src[i]) return VSE_ERR_ARG;
  const int np = (int)s->axis_len(axis) - 2 * (s->d.ord - 1);
  // active band: a != 0 or k != 1 on either grid
  std::vector<char> act(np);
  for (int j = 0; j < np; j++) act[j] = a[j] != 0.f || k[j] != 1.f || a_h[j] != 0.f || k_h[j] != 1.f;
  int lo = 0;
  while (lo < np && act[lo]) lo++;
  int hi = np;
  while (hi > lo && act[hi - 1]) hi--;
  for (int j = lo; j < hi; j++)
    if (act[j]) { lo = np; hi = np; break; }   // not two end bands: keep the whole axis
  if (s->psi_ready) {
    AxisProf &o = s->ax[axis];
    if (o.lo != lo || o.hi != hi) return s->fail(VSE_ERR_STATE, "CPML band changed after state allocation");
  }
  int rc = 0;
  std::vector<float> inv_k(np), inv_kh(np);
  SyncOnExit sync(s->stream);     // declared after the vectors: waits before they are freed
  for (int j = 0; j < np; j++) {
    inv_k[j] = 1.0f / k[j];
    inv_kh[j] = 1.0f / k_h[j];
  }
  const float *all[8] = {a, b, k, a_h, b_h, k_h, inv_k.data(), inv_kh.data()};
  for (int i = 0; i < 8 && !rc; i++) {
    if (!s->prof[axis][i]) rc = s->alloc(&s->prof[axis][i], np);
    if (!rc) rc = s->hip(hipMemcpyAsync(s->prof[axis][i], all[i], np * sizeof(float), hipMemcpyHostToDevice,
                                        s->stream), "copy cpml");
  }
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "set_cpml copy");
  if (rc) return rc;
  AxisProf &P = s->ax[axis];
  P.a = s->prof[axis][0]; P.b = s->prof[axis][1]; P.k = s->prof[axis][2];
  P.ah = s->prof[axis][3]; P.bh = s->prof[axis][4]; P.kh = s->prof[axis][5];
  P.ki = s->prof[axis][6]; P.khi = s->prof[axis][7];
  P.lo = lo; P.hi = hi; P.np = np; P.nb = lo + (np - hi);
  s->have_cpml[axis] = true;
  s->drop_graph();
  return s->hip(hipStreamSynchronize(s->stream), "set_cpml");
}

int vse_set_sources(vse_sim *s, int n_pts, const int *cells, const int *cols, int n_cols, const float *src) {
  if (check_handle(s) || n_pts < 0 || n_cols < 1 || !src || (n_pts > 0 && (!cells || !cols))) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  for (int i = 0; i < n_pts; i++)
    if (cells[i] < 0 || cells[i] >= s->ncell || cols[i] < 0 || cols[i] >= n_cols)
      return s->fail(VSE_ERR_ARG, "source point out of range");
  {
    // one column per cell, as the reference's per-cell source map
    std::vector<int> sorted(cells, cells + n_pts);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return s->fail(VSE_ERR_ARG, "source cell listed twice");
  }
  std::vector<int> map, rows, rbeg, rz, rcol;
  SyncOnExit sync(s->stream);     // after the vectors: waits before they are freed
  // the cached graph holds the old buffers' addresses: drop it before they go
  s->drop_graph();
  (void)hipStreamSynchronize(s->stream);
  s->release(s->src_cell); s->release(s->src_col); s->release(s->src); s->release(s->src_map);
  s->release(s->src_rowflag);
  s->release(s->src_rbeg); s->release(s->src_rz); s->release(s->src_rcol);
  s->src_cell = s->src_col = s->src_map = s->src_rowflag = nullptr;
  s->src_rbeg = s->src_rz = s->src_rcol = nullptr;
  s->src = nullptr;
  s->n_src_pts = 0;
  int rc = s->alloc(&s->src_cell, n_pts);
  if (!rc) rc = s->alloc(&s->src_col, n_pts);
  if (!rc) rc = s->alloc(&s->src, (size_t)s->d.n_rows * n_cols);
  if (!rc && n_pts) rc = s->hip(hipMemcpyAsync(s->src_cell, cells, n_pts * sizeof(int), hipMemcpyHostToDevice, s->stream), "src cells");
  if (!rc && n_pts) rc = s->hip(hipMemcpyAsync(s->src_col, cols, n_pts * sizeof(int), hipMemcpyHostToDevice, s->stream), "src cols");
  if (!rc) rc = s->hip(hipMemcpyAsync(s->src, src, (size_t)s->d.n_rows * n_cols * sizeof(float), hipMemcpyHostToDevice, s->stream), "src");
  if (!rc && s->d.ndim == 2) {     // 2D velocity kernels inject through a per-cell map
    map.assign(s->ncell, -1);
    rows.assign(s->d.nx, 0);
    for (int i = 0; i < n_pts; i++) {
      map[cells[i]] = cols[i];
      rows[cells[i] / s->d.ny] = 1;
    }
    rc = s->alloc(&s->src_map, s->ncell);
    if (!rc) rc = s->hip(hipMemcpyAsync(s->src_map, map.data(), s->ncell * sizeof(int), hipMemcpyHostToDevice, s->stream), "src map");
    if (!rc) rc = s->alloc(&s->src_rowflag, s->d.nx);
    if (!rc) rc = s->hip(hipMemcpyAsync(s->src_rowflag, rows.data(), s->d.nx * sizeof(int), hipMemcpyHostToDevice, s->stream), "src rows");
  }
  if (!rc && s->d.ndim == 3) {     // the fused 3D step finds its sources by (x, y) row
    const long nrow = (long)s->d.nx * s->d.ny;
    rbeg.assign(nrow + 1, 0);
    for (int i = 0; i < n_pts; i++) rbeg[cells[i] / s->d.nz + 1]++;
    for (long r = 0; r < nrow; r++) rbeg[r + 1] += rbeg[r];
    rz.resize(std::max(n_pts, 1));
    rcol.resize(std::max(n_pts, 1));
    std::vector<int> fill(rbeg.begin(), rbeg.end() - 1);
    for (int i = 0; i < n_pts; i++) {
      const int q = fill[cells[i] / s->d.nz]++;
      rz[q] = cells[i] % s->d.nz;
      rcol[q] = cols[i];
    }
    rows.assign(s->d.nx, 0);       // per x-plane: 1 if the plane holds a source cell
    for (int i = 0; i < n_pts; i++) rows[cells[i] / ((long)s->d.ny * s->d.nz)] = 1;
    rc = s->alloc(&s->src_rowflag, s->d.nx);
    if (!rc) rc = s->hip(hipMemcpyAsync(s->src_rowflag, rows.data(), s->d.nx * sizeof(int), hipMemcpyHostToDevice, s->stream), "src planes");
    if (!rc) rc = s->alloc(&s->src_rbeg, nrow + 1);
    if (!rc) rc = s->alloc(&s->src_rz, n_pts);
    if (!rc) rc = s->alloc(&s->src_rcol, n_pts);
    if (!rc) rc = s->hip(hipMemcpyAsync(s->src_rbeg, rbeg.data(), (nrow + 1) * sizeof(int), hipMemcpyHostToDevice, s->stream), "src row table");
    if (!rc && n_pts) rc = s->hip(hipMemcpyAsync(s->src_rz, rz.data(), n_pts * sizeof(int), hipMemcpyHostToDevice, s->stream), "src row z");
    if (!rc && n_pts) rc = s->hip(hipMemcpyAsync(s->src_rcol, rcol.data(), n_pts * sizeof(int), hipMemcpyHostToDevice, s->stream), "src row col");
  }
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "set_sources");
  if (rc) {
    // leave no partial source set: vse_run then refuses the simulation instead
    // of launching kernels on a missing map or row table
    (void)hipStreamSynchronize(s->stream);
    s->release(s->src_cell); s->release(s->src_col); s->release(s->src); s->release(s->src_map);
    s->release(s->src_rowflag);
    s->release(s->src_rbeg); s->release(s->src_rz); s->release(s->src_rcol);
    s->src_cell = s->src_col = s->src_map = s->src_rowflag = nullptr;
    s->src_rbeg = s->src_rz = s->src_rcol = nullptr;
    s->src = nullptr;
    s->n_src_pts = 0;
    return rc;
  }
  s->n_src_pts = n_pts;
  s->src_cols = n_cols;
  s->drop_graph();
  return VSE_OK;
}

int vse_set_receivers(vse_sim *s, int n_rec, const int *beg, const int *cells, const int *delay, int gate,
                      int channel) {
  if (check_handle(s) || n_rec < 0 || (n_rec > 0 && (!beg || !delay))) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (channel < 0) channel = s->d.ndim == 2 ? VSE_F_SYY : VSE_F_SZZ;
  if (!field_present(s->d, channel)) return s->fail(VSE_ERR_ARG, "receiver channel is not a field of this simulation");
  const int n_pts = n_rec > 0 ? beg[n_rec] : 0;
  if (n_rec > 0 && beg[0] != 0) return s->fail(VSE_ERR_ARG, "receiver CSR must start at 0");
  for (int r = 0; r < n_rec; r++)
    if (beg[r + 1] < beg[r]) return s->fail(VSE_ERR_ARG, "receiver CSR not monotone");
  if (n_pts > 0 && !cells) return VSE_ERR_ARG;
  for (int p = 0; p < n_pts; p++)
    if (cells[p] < 0 || cells[p] >= s->ncell) return s->fail(VSE_ERR_ARG, "receiver point out of range");
  SyncOnExit sync(s->stream);
  // the cached graph holds the old buffers' addresses: drop it before they go
  s->drop_graph();
  (void)hipStreamSynchronize(s->stream);
  s->release(s->rec_beg); s->release(s->rec_cell); s->release(s->rec_delay); s->release(s->traces);
  s->rec_beg = s->rec_cell = s->rec_delay = nullptr;
  s->traces = nullptr;
  s->n_rec = 0;
  int rc = s->alloc(&s->rec_beg, n_rec + 1);
  if (!rc) rc = s->alloc(&s->rec_cell, n_pts);
  if (!rc) rc = s->alloc(&s->rec_delay, n_rec);
  if (!rc) rc = s->alloc(&s->traces, (size_t)s->d.n_rows * std::max(n_rec, 1));
  if (!rc && n_rec) rc = s->hip(hipMemcpyAsync(s->rec_beg, beg, (n_rec + 1) * sizeof(int), hipMemcpyHostToDevice, s->stream), "rec beg");
  if (!rc && n_pts) rc = s->hip(hipMemcpyAsync(s->rec_cell, cells, n_pts * sizeof(int), hipMemcpyHostToDevice, s->stream), "rec cells");
  if (!rc && n_rec) rc = s->hip(hipMemcpyAsync(s->rec_delay, delay, n_rec * sizeof(int), hipMemcpyHostToDevice, s->stream), "rec delay");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "set_receivers");
  if (rc) return rc;
  s->n_rec = n_rec;
  s->rec_gate = gate;
  s->rec_channel = channel;
  return VSE_OK;
}

static int finalize(vse_sim *s) {
  if (!s->have_material) return s->fail(VSE_ERR_STATE, "vse_set_material not called");
  for (int a = 0; a < s->d.ndim; a++)
    if (!s->have_cpml[a]) return s->fail(VSE_ERR_STATE, "vse_set_cpml not called for every axis");
  if (!s->src) return s->fail(VSE_ERR_STATE, "vse_set_sources not called");
  if (!s->rec_beg) return s->fail(VSE_ERR_STATE, "vse_set_receivers not called");
  if (s->psi_ready) return 0;
  for (int slot = 0; slot < P_COUNT; slot++) {
    if (!s->slot_used(slot)) continue;
    int rc = s->alloc(&s->psi[slot], s->psi_elems(slot));
    if (rc) return rc;
  }
  // every slot the kernels index must exist before any launch
  for (int slot = 0; slot < P_COUNT; slot++)
    if (s->slot_used(slot) && !s->psi[slot]) return s->fail(VSE_ERR_STATE, "psi slot not allocated");
  for (int a = 0; a < s->d.ndim; a++)
    for (int i = 0; i < 8; i++)
      if (!s->prof[a][i]) return s->fail(VSE_ERR_STATE, "CPML profile not allocated");
  for (int f = 0; f < 9; f++) {
    const bool need = s->d.ndim == 3 || f == VSE_F_VX || f == VSE_F_VY || f == VSE_F_SXX || f == VSE_F_SYY ||
                      f == VSE_F_SXY;
    if (need && !s->field[f]) return s->fail(VSE_ERR_STATE, "field not allocated");
  }
  if (s->d.visco && s->d.variant == VSE_VARIANT_WGSL)
    for (int c = 0; c < 6; c++)
      if ((s->d.ndim == 3 || c == 0 || c == 1 || c == 3) && !s->r[c])
        return s->fail(VSE_ERR_STATE, "memory variable not allocated");
  s->psi_ready = true;
  return s->hip(hipStreamSynchronize(s->stream), "finalize");
}

int vse_reset(vse_sim *s) {
  if (check_handle(s)) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  int rc = finalize(s);
  if (rc) return rc;
  for (int f = 0; f < 9 && !rc; f++)
    if (s->field[f]) rc = s->hip(hipMemsetAsync(s->field[f], 0, s->ncell * sizeof(float), s->stream), "reset");
  for (int c = 0; c < 6 && !rc; c++)
    if (s->r[c]) rc = s->hip(hipMemsetAsync(s->r[c], 0, (size_t)s->rs * s->d.n_sls * sizeof(float), s->stream), "reset r");
  for (int p = 0; p < P_COUNT && !rc; p++)
    if (s->psi[p]) rc = s->hip(hipMemsetAsync(s->psi[p], 0, std::max<long>(s->psi_elems(p), 1) * sizeof(float), s->stream), "reset psi");
  if (!rc) rc = s->hip(hipMemsetAsync(s->traces, 0, (size_t)s->d.n_rows * std::max(s->n_rec, 1) * sizeof(float), s->stream), "reset traces");
  if (!rc) rc = s->hip(hipMemsetAsync(s->ctrl, 0, 2 * sizeof(unsigned), s->stream), "reset ctrl");
  if (!rc && s->v2slots)
    rc = s->hip(hipMemsetAsync(s->v2slots, 0, (size_t)s->n_slots * sizeof(unsigned), s->stream), "reset slots");
  for (int f = 0; f < 9 && !rc && s->alt_ready; f++)
    if (s->alt[f]) rc = s->hip(hipMemsetAsync(s->alt[f], 0, s->ncell * sizeof(float), s->stream), "reset");
  for (int p = P_DSXX_DX; p < P_COUNT && !rc && s->alt_ready; p++)
    if (s->psi_alt[p])
      rc = s->hip(hipMemsetAsync(s->psi_alt[p], 0, std::max<long>(s->psi_elems(p), 1) * sizeof(float), s->stream),
                  "reset psi");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "reset");
  return rc;
}

static void make_args(vse_sim *s, StepArgs &A, SparseArgs &S) {
  const vse_desc &d = s->d;
  memset(&A, 0, sizeof(A));
  A.nx = d.nx; A.ny = d.ny; A.nz = d.nz; A.ncell = s->ncell;
  A.dx = d.dx; A.dy = d.dy; A.dz = d.ndim == 3 ? d.dz : 1.f; A.dt = d.dt;
  A.one_dx = 1.0f / d.dx; A.one_dy = 1.0f / d.dy; A.one_dz = d.ndim == 3 ? 1.0f / d.dz : 1.f;
  for (int c = 0; c < d.ord; c++) {
    A.C[c] = d.coefs[c];
    A.Cd[0][c] = d.coefs[c] / A.dx;
    A.Cd[1][c] = d.coefs[c] / A.dy;
    A.Cd[2][c] = d.coefs[c] / A.dz;
  }
  for (int a = 0; a < 3; a++) A.ax[a] = s->ax[a];
  A.vx = s->field[VSE_F_VX]; A.vy = s->field[VSE_F_VY]; A.vz = s->field[VSE_F_VZ];
  A.sxx = s->field[VSE_F_SXX]; A.syy = s->field[VSE_F_SYY]; A.szz = s->field[VSE_F_SZZ];
  A.sxy = s->field[VSE_F_SXY]; A.sxz = s->field[VSE_F_SXZ]; A.syz = s->field[VSE_F_SYZ];
  for (int p = 0; p < P_COUNT; p++) A.psi[p] = s->psi[p];
  for (int c = 0; c < 6; c++) A.r[c] = s->r[c];
  A.rs = s->rs;
  A.rho = s->rho; A.cp = s->cp; A.cs = s->cs;
  A.zk = s->zk; A.L = d.n_sls;
  A.v2max = s->ctrl;
  A.v2slots = s->v2slots; A.n_slots = s->n_slots;
  A.step = (const int *)(s->ctrl + 1);
  A.src_map = s->src_map; A.src = s->src; A.src_cols = s->src_cols; A.src_rows = d.n_rows;
  A.src_rowflag = s->src_rowflag;
  A.n_rec = s->n_rec; A.rec_gate = s->rec_gate; A.trace_rows = d.n_rows;
  A.rec_beg = s->rec_beg; A.rec_cell = s->rec_cell; A.rec_delay = s->rec_delay; A.traces = s->traces;
  A.rec_field = s->field[s->rec_channel];
  A.rec_after = velocity_field(s->rec_channel) ? 1 : 0;
  A.src_rbeg = s->src_rbeg; A.src_rz = s->src_rz; A.src_rcol = s->src_rcol;

  memset(&S, 0, sizeof(S));
  S.ndim = d.ndim; S.nx = d.nx; S.ny = d.ny; S.nz = d.nz; S.dt = d.dt;
  S.vx = A.vx; S.vy = A.vy; S.vz = A.vz;
  S.rec_field = s->field[s->rec_channel];
  S.rho = s->rho;
  S.v2max = s->ctrl;
  S.step = (int *)(s->ctrl + 1);
  S.n_src_pts = d.variant == VSE_VARIANT_TAICHI ? 0 : s->n_src_pts;
  S.src_cell = s->src_cell; S.src_col = s->src_col; S.src = s->src; S.src_cols = s->src_cols; S.src_rows = d.n_rows;
  S.n_rec = s->n_rec; S.rec_beg = s->rec_beg; S.rec_cell = s->rec_cell; S.rec_delay = s->rec_delay;
  S.rec_gate = s->rec_gate; S.traces = s->traces; S.trace_rows = d.n_rows;
}

// Step i of a run of k launched together.  2D: the velocity kernel advances
// step *counter + i and does its sources and receivers, and advance_steps()
// moves the counter after the k steps.  3D: the sparse kernel does that work
// and increments the counter itself.
// The counter, running max and slots are the ones A and S point at (the
// handle's own, or the no-allocation entry point's scratch); st = 0 is the
// handle's stream.
static hipError_t enqueue_step(vse_sim *s, const StepArgs &A, const SparseArgs &S, int i, hipStream_t st = 0) {
  const vse_desc &d = s->d;
  if (!st) st = s->stream;
  hipError_t e;
  if (d.ndim == 2) {
    e = vse_launch_stress2d(d.variant, d.visco, d.ord, s->math, A, s->tune[0], st);
    if (e != hipSuccess) return e;
    e = vse_launch_velocity2d(d.variant, d.visco, d.ord, s->math, A, s->tune[1], i, st);
    if (e == hipSuccess && A.rec_after) e = vse_launch_record2d(A, i, st);
    return e;
  } else {
    e = vse_launch_stress3d(d.visco, d.ord, s->math, A, s->tune[0], st);
    if (e != hipSuccess) return e;
    e = vse_launch_velocity3d(d.visco, d.ord, s->math, A, s->tune[1], st);
  }
  if (e != hipSuccess) return e;
  return vse_launch_sparse(S, st);
}

static hipError_t advance_steps(vse_sim *s, const StepArgs &A, int k, hipStream_t st = 0) {
  return s->d.ndim == 2 ? vse_launch_advance((int *)A.step, k, st ? st : s->stream) : hipSuccess;
}

static int get_graph(vse_sim *s, int steps, const StepArgs &A, const SparseArgs &S) {
  if (s->gexec && s->gexec_steps == steps) return 0;
  s->drop_graph();
  hipGraph_t g = nullptr;
  int rc = s->hip(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  hipError_t e = hipSuccess;
  for (int i = 0; i < steps && e == hipSuccess; i++) e = enqueue_step(s, A, S, i);
  if (e == hipSuccess) e = advance_steps(s, A, steps);
  hipError_t e2 = hipStreamEndCapture(s->stream, &g);
  if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return s->hip(e, "capture launch"); }
  rc = s->hip(e2, "end capture");
  if (rc) return rc;
  rc = s->hip(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(g);
  if (!rc) s->gexec_steps = steps;
  return rc;
}

// ctrl[0] = max(ctrl[0], every workgroup slot) (2D; 3D kernels update ctrl[0])
static int fold_v2(vse_sim *s) {
  return s->v2slots ? s->hip(vse_launch_fold_v2(s->v2slots, s->n_slots, s->ctrl, s->stream), "fold v2") : 0;
}

static int check_stable(vse_sim *s) {
  int rc = fold_v2(s);
  if (!rc) rc = s->hip(hipMemcpyAsync(s->host_ctrl, s->ctrl, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream), "read v2max");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "sync");
  if (rc) return rc;
  if (s->host_ctrl[0] >= 0x7f800000u) {
    float v;
    memcpy(&v, &s->host_ctrl[0], 4);
    char buf[128];
    snprintf(buf, sizeof buf, "Simulacao tornando-se instavel [|v|^2 = %g at step %u]", v, s->host_ctrl[1]);
    return s->fail(VSE_ERR_UNSTABLE, buf);
  }
  return 0;
}

// ---- fused 3D schedule --------------------------------------------------------
// A run of n steps from state k: the stress pass of step k in place on set 0;
// n-1 fused steps, each reading one buffer set and writing the other
// (velocity of step m, stress of step m+1), each followed by the sparse kernel
// of step m (receivers, counter); the velocity pass and sparse kernel of the
// last step in place on the set in use.  The handle's field[] / psi[] then
// point at the final state (the two sets are swapped when it ends in set 1).

// The 2D fused kernel keeps one |v|^2 slot per workgroup: its grid must fit
// the slots the two-pass velocity launch allocated.
static bool fused2d_fits(const vse_sim *s, int xchunk) {
  return s->d.ndim == 2 && xchunk >= 1 && vse_fused2d_blocks(s->d.nx, s->d.ny, s->d.ord, xchunk) <= s->n_slots;
}

static bool fused_active(const vse_sim *s) {
  if (s->sched != VSE_SCHED_FUSED) return false;
  if (s->d.ndim == 2) return fused2d_fits(s, s->tune[2].xchunk);
  return s->d.variant == VSE_VARIANT_WGSL && vse_fused3d_supported(s->ncell, s->rs, s->d.n_sls, s->d.visco, s->d.ord);
}

static float *set_field(vse_sim *s, int set, int f) { return set ? s->alt[f] : s->field[f]; }
static float *set_psi(vse_sim *s, int set, int p) { return set && p >= P_DSXX_DX ? s->psi_alt[p] : s->psi[p]; }

// The second buffer set (sigma, v, velocity-pass psi), zeroed, on first use.
static int ensure_alt(vse_sim *s) {
  if (s->alt_ready) return 0;
  int rc = 0;
  for (int f = 0; f < 9 && !rc; f++)
    if (s->field[f]) rc = s->alloc(&s->alt[f], s->ncell);
  for (int p = P_DSXX_DX; p < P_COUNT && !rc; p++)
    if (s->psi[p]) rc = s->alloc(&s->psi_alt[p], s->psi_elems(p));
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "second buffer set");
  if (rc) {
    for (int f = 0; f < 9; f++) { s->release(s->alt[f]); s->alt[f] = nullptr; }
    for (int p = 0; p < P_COUNT; p++) { s->release(s->psi_alt[p]); s->psi_alt[p] = nullptr; }
    return rc;
  }
  s->alt_ready = true;
  return 0;
}

// make_args() of the state held in buffer set `set`
static void make_args_set(vse_sim *s, StepArgs &A, SparseArgs &S, int set) {
  make_args(s, A, S);
  if (!set) return;
  A.vx = s->alt[VSE_F_VX]; A.vy = s->alt[VSE_F_VY]; A.vz = s->alt[VSE_F_VZ];
  A.sxx = s->alt[VSE_F_SXX]; A.syy = s->alt[VSE_F_SYY]; A.szz = s->alt[VSE_F_SZZ];
  A.sxy = s->alt[VSE_F_SXY]; A.sxz = s->alt[VSE_F_SXZ]; A.syz = s->alt[VSE_F_SYZ];
  for (int p = P_DSXX_DX; p < P_COUNT; p++) A.psi[p] = s->psi_alt[p];
  A.rec_field = s->alt[s->rec_channel];
  S.vx = A.vx; S.vy = A.vy; S.vz = A.vz;
  S.rec_field = A.rec_field;
}

// One fused step reading set p and writing set 1-p, then the sparse kernel of
// that step without sources (the fused kernel injected them): a stress channel
// is read from set p (step m+1's stress is in set 1-p already), a velocity
// channel from set 1-p.
// 2D: the fused launch advances step *step + i (the counter moves once per
// graph, as in the two-pass loop) and records a stress channel itself; a
// velocity channel is gathered from set 1-p after it.
static hipError_t enqueue_fused(vse_sim *s, int p, hipStream_t st, hipEvent_t mid = nullptr, int i = 0) {
  FusedArgs F;
  SparseArgs S;
  make_args_set(s, F.A, S, p);
  const int o = 1 - p;
  const int sig[6] = {VSE_F_SXX, VSE_F_SYY, VSE_F_SZZ, VSE_F_SXY, VSE_F_SXZ, VSE_F_SYZ};
  for (int c = 0; c < 3; c++) F.vo[c] = set_field(s, o, VSE_F_VX + c);
  for (int c = 0; c < 6; c++) F.so[c] = set_field(s, o, sig[c]);
  for (int k = 0; k < 9; k++) F.psio[k] = set_psi(s, o, P_DSXX_DX + k);
  if (s->d.ndim == 2) {
    hipError_t e = vse_launch_fused2d(s->d.variant, s->d.visco, s->d.ord, s->math, F, s->tune[2].xchunk, i, st);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
    if (e == hipSuccess && F.A.rec_after) {
      StepArgs B;
      make_args_set(s, B, S, o);
      e = vse_launch_record2d(B, i, st);
    }
    return e;
  }
  hipError_t e = vse_launch_fused3d(s->math, F, s->tune[2], st);
  if (e == hipSuccess && mid) e = hipEventRecord(mid, st);
  if (e != hipSuccess) return e;
  S.n_src_pts = 0;
  S.rec_field = set_field(s, velocity_field(s->rec_channel) ? o : p, s->rec_channel);
  return vse_launch_sparse(S, st);
}

// graph of `steps` (even) fused steps starting from set 0
static int get_fused_graph(vse_sim *s, int steps) {
  if (s->fgexec && s->fgexec_steps == steps) return 0;
  if (s->fgexec) (void)hipGraphExecDestroy(s->fgexec);
  s->fgexec = nullptr;
  s->fgexec_steps = 0;
  hipGraph_t g = nullptr;
  int rc = s->hip(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  hipError_t e = hipSuccess;
  for (int i = 0; i < steps && e == hipSuccess; i++) e = enqueue_fused(s, i & 1, s->stream, nullptr, i);
  if (e == hipSuccess && s->d.ndim == 2) e = vse_launch_advance(s->ctrl_step(), steps, s->stream);
  hipError_t e2 = hipStreamEndCapture(s->stream, &g);
  if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return s->hip(e, "capture launch"); }
  rc = s->hip(e2, "end capture");
  if (rc) return rc;
  rc = s->hip(hipGraphInstantiate(&s->fgexec, g, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(g);
  if (!rc) s->fgexec_steps = steps;
  return rc;
}

// the final state is in set 1: make field[] / psi[] point at it
static void swap_sets(vse_sim *s) {
  for (int f = 0; f < 9; f++) std::swap(s->field[f], s->alt[f]);
  for (int p = P_DSXX_DX; p < P_COUNT; p++) std::swap(s->psi[p], s->psi_alt[p]);
  s->drop_graph();
}

// The last step's velocity pass and sparse kernel, in place on set p.
static hipError_t enqueue_fused_tail(vse_sim *s, int p) {
  StepArgs A;
  SparseArgs S;
  make_args_set(s, A, S, p);
  if (s->d.ndim == 2) {
    hipError_t e = vse_launch_velocity2d(s->d.variant, s->d.visco, s->d.ord, s->math, A, s->tune[1], 0, s->stream);
    if (e == hipSuccess && A.rec_after) e = vse_launch_record2d(A, 0, s->stream);
    return e == hipSuccess ? vse_launch_advance(s->ctrl_step(), 1, s->stream) : e;
  }
  hipError_t e = vse_launch_velocity3d(s->d.visco, s->d.ord, s->math, A, s->tune[1], s->stream);
  return e == hipSuccess ? vse_launch_sparse(S, s->stream) : e;
}

// The stress pass that opens a fused run, in place on set 0.
static hipError_t enqueue_fused_head(vse_sim *s, const StepArgs &A) {
  if (s->d.ndim == 2) return vse_launch_stress2d(s->d.variant, s->d.visco, s->d.ord, s->math, A, s->tune[0], s->stream);
  return vse_launch_stress3d(s->d.visco, s->d.ord, s->math, A, s->tune[0], s->stream);
}

static int run_fused(vse_sim *s, int n_steps, int check_every, float *elapsed_ms) {
  int rc = ensure_alt(s);
  if (rc) return rc;
  StepArgs A;
  SparseArgs S;
  make_args(s, A, S);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    rc = s->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = s->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = s->hip(hipEventRecord(e0, s->stream), "event record");
  }
  if (!rc) rc = s->hip(enqueue_fused_head(s, A), "stress launch");
  const int nf = n_steps - 1;
  int chunk = std::min(std::min(check_every > 0 ? check_every : 256, 256), nf) & ~1;
  if (!rc && chunk >= 2) rc = get_fused_graph(s, chunk);
  int p = 0, done = 0, since_check = 0;
  while (!rc && done < nf) {
    const int left = nf - done;
    if (chunk >= 2 && left >= chunk) {
      rc = s->hip(hipGraphLaunch(s->fgexec, s->stream), "graph launch");
      done += chunk;
      since_check += chunk;
    } else {
      for (int i = 0; i < left && !rc; i++, p ^= 1) rc = s->hip(enqueue_fused(s, p, s->stream, nullptr, i), "fused launch");
      if (!rc && s->d.ndim == 2) rc = s->hip(vse_launch_advance(s->ctrl_step(), left, s->stream), "step launch");
      done += left;
      since_check += left;
    }
    if (!rc && check_every > 0 && since_check >= check_every && done < nf) {
      rc = check_stable(s);
      since_check = 0;
    }
  }
  if (!rc) rc = s->hip(enqueue_fused_tail(s, p), "velocity launch");
  if (elapsed_ms && !rc) {
    rc = s->hip(hipEventRecord(e1, s->stream), "event record");
    if (!rc) rc = s->hip(hipEventSynchronize(e1), "event sync");
    if (!rc) rc = s->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc && p == 1) {
    (void)hipStreamSynchronize(s->stream);
    swap_sets(s);
  }
  if (!rc) rc = check_stable(s);
  return rc;
}

// vse_profile of the fused schedule: n_steps steps as run_fused does them,
// eagerly; ms[0] = average fused kernel, ms[1] = 0, ms[2] = average sparse kernel.
static int profile_fused(vse_sim *s, int n_steps, float *ms) {
  if (n_steps < 2) return s->fail(VSE_ERR_ARG, "profiling the fused schedule takes 2 steps or more");
  int rc = ensure_alt(s);
  if (rc) return rc;
  StepArgs A;
  SparseArgs S;
  make_args(s, A, S);
  hipEvent_t ev[3] = {};
  for (int i = 0; i < 3 && !rc; i++) rc = s->hip(hipEventCreate(&ev[i]), "event");
  if (!rc) rc = s->hip(enqueue_fused_head(s, A), "stress launch");
  double acc[2] = {0, 0};
  int p = 0;
  for (int it = 0; it < n_steps - 1 && !rc; it++, p ^= 1) {
    hipError_t e = hipEventRecord(ev[0], s->stream);
    if (e == hipSuccess) e = enqueue_fused(s, p, s->stream, ev[1]);
    if (e == hipSuccess && s->d.ndim == 2) e = vse_launch_advance(s->ctrl_step(), 1, s->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[2], s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
    rc = s->hip(e, "profile step");
    for (int k = 0; k < 2 && !rc; k++) {
      float t = 0.f;
      rc = s->hip(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "elapsed");
      acc[k] += t;
    }
  }
  if (!rc) rc = s->hip(enqueue_fused_tail(s, p), "velocity launch");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "profile");
  for (int i = 0; i < 3; i++)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (!rc && p == 1) swap_sets(s);
  if (!rc) {
    ms[0] = (float)(acc[0] / (n_steps - 1));
    ms[1] = 0.f;
    ms[2] = (float)(acc[1] / (n_steps - 1));
    rc = check_stable(s);
  }
  return rc;
}

int vse_set_schedule(vse_sim *s, int sched) {
  if (check_handle(s) || (sched != VSE_SCHED_TWO_PASS && sched != VSE_SCHED_FUSED)) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (sched == VSE_SCHED_FUSED && s->d.ndim == 2 && !fused2d_fits(s, s->tune[2].xchunk))
    return s->fail(VSE_ERR_ARG, "the fused 2D step needs x-chunks of at least ~4 rows (one |v|^2 slot per workgroup)");
  if (sched == VSE_SCHED_FUSED && s->d.ndim == 3 &&
      !(s->d.variant == VSE_VARIANT_WGSL && vse_fused3d_supported(s->ncell, s->rs, s->d.n_sls, s->d.visco, s->d.ord)))
    return s->fail(VSE_ERR_ARG, "the fused step covers 3D viscoelastic simulations with 3 mechanisms and FD order 2 "
                                "(acc = 2), with 3*cells*4 bytes below 2^32");
  s->sched = sched;
  s->drop_graph();
  return VSE_OK;
}

int vse_run(vse_sim *s, int n_steps, int check_every, float *elapsed_ms) {
  if (check_handle(s) || n_steps < 0) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  int rc = finalize(s);
  if (rc) return rc;
  if (n_steps == 0) { if (elapsed_ms) *elapsed_ms = 0.f; return 0; }
  if (fused_active(s)) return run_fused(s, n_steps, check_every, elapsed_ms);
  StepArgs A;
  SparseArgs S;
  make_args(s, A, S);
  const int chunk = std::max(1, std::min(check_every > 0 ? check_every : 256, std::min(n_steps, 256)));
  rc = get_graph(s, chunk, A, S);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    rc = s->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = s->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = s->hip(hipEventRecord(e0, s->stream), "event record");
  }
  int done = 0, since_check = 0;
  while (!rc && done < n_steps) {
    const int left = n_steps - done;
    if (left >= chunk) {
      rc = s->hip(hipGraphLaunch(s->gexec, s->stream), "graph launch");
      done += chunk;
      since_check += chunk;
    } else {
      for (int i = 0; i < left && !rc; i++) rc = s->hip(enqueue_step(s, A, S, i), "step launch");
      if (!rc) rc = s->hip(advance_steps(s, A, left), "step launch");
      done += left;
      since_check += left;
    }
    if (!rc && check_every > 0 && since_check >= check_every && done < n_steps) {
      rc = check_stable(s);
      since_check = 0;
    }
  }
  if (elapsed_ms && !rc) {
    rc = s->hip(hipEventRecord(e1, s->stream), "event record");
    if (!rc) rc = s->hip(hipEventSynchronize(e1), "event sync");
    if (!rc) rc = s->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = check_stable(s);
  return rc;
}

int vse_run_noalloc(vse_sim *s, int n_steps, void *stream) {
  if (check_handle(s) || n_steps < 0) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (!s->psi_ready || !s->have_material || !s->src || !s->rec_beg)
    return s->fail(VSE_ERR_STATE, "vse_run_noalloc: call vse_reset (or vse_run) first; that allocates the state");
  if (s->n_slots > VSE_NOALLOC_SLOTS)
    return s->fail(VSE_ERR_ARG, "vse_run_noalloc: the grid has " + std::to_string(s->n_slots) +
                                    " velocity workgroups, the scratch holds VSE_NOALLOC_SLOTS = " +
                                    std::to_string(VSE_NOALLOC_SLOTS));
  unsigned *scr = scratch_of(s->device);
  if (!scr) return s->fail(VSE_ERR_STATE, "vse_run_noalloc: no scratch address for this device");
  if (n_steps == 0) return 0;
  const hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  StepArgs A;
  SparseArgs S;
  make_args(s, A, S);
  A.v2max = scr;
  A.step = (const int *)(scr + 1);
  A.v2slots = s->n_slots ? scr + 2 : nullptr;
  S.v2max = scr;
  S.step = (int *)(scr + 1);
  const int words = 2 + s->n_slots;
  hipLaunchKernelGGL(k_scratch_begin, dim3((words + 255) / 256), dim3(256), 0, st, scr, s->ctrl, s->n_slots);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < n_steps && e == hipSuccess; i++) e = enqueue_step(s, A, S, i, st);
  if (e == hipSuccess) e = advance_steps(s, A, n_steps, st);
  if (e == hipSuccess && s->n_slots) e = vse_launch_fold_v2(scr + 2, s->n_slots, scr, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_scratch_end, dim3(1), dim3(64), 0, st, scr, s->ctrl);
    e = hipGetLastError();
  }
  return s->hip(e, "vse_run_noalloc launch");
}

int vse_profile(vse_sim *s, int n_steps, float *ms) {
  if (check_handle(s) || n_steps < 1 || !ms) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  int rc = finalize(s);
  if (rc) return rc;
  if (fused_active(s)) return profile_fused(s, n_steps, ms);
  StepArgs A;
  SparseArgs S;
  make_args(s, A, S);
  const vse_desc &d = s->d;
  hipEvent_t ev[4] = {};
  for (int i = 0; i < 4 && !rc; i++) rc = s->hip(hipEventCreate(&ev[i]), "event");
  double acc[3] = {0, 0, 0};
  for (int it = 0; it < n_steps && !rc; it++) {
    hipError_t e = hipEventRecord(ev[0], s->stream);
    if (e == hipSuccess) e = d.ndim == 2 ? vse_launch_stress2d(d.variant, d.visco, d.ord, s->math, A, s->tune[0], s->stream)
                                         : vse_launch_stress3d(d.visco, d.ord, s->math, A, s->tune[0], s->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[1], s->stream);
    if (e == hipSuccess) e = d.ndim == 2 ? vse_launch_velocity2d(d.variant, d.visco, d.ord, s->math, A, s->tune[1], 0, s->stream)
                                         : vse_launch_velocity3d(d.visco, d.ord, s->math, A, s->tune[1], s->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[2], s->stream);
    if (e == hipSuccess && d.ndim == 2 && A.rec_after) e = vse_launch_record2d(A, 0, s->stream);
    if (e == hipSuccess) e = d.ndim == 2 ? advance_steps(s, A, 1) : vse_launch_sparse(S, s->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[3], s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[3]);
    rc = s->hip(e, "profile step");
    for (int k = 0; k < 3 && !rc; k++) {
      float t = 0.f;
      rc = s->hip(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "elapsed");
      acc[k] += t;
    }
  }
  for (int i = 0; i < 4; i++)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (!rc)
    for (int k = 0; k < 3; k++) ms[k] = (float)(acc[k] / n_steps);
  if (!rc) rc = check_stable(s);
  return rc;
}

int vse_set_tuning(vse_sim *s, int kernel, int ty, int nt, int xchunk, int remap) {
  if (check_handle(s) || kernel < 0 || kernel > 2) return VSE_ERR_ARG;
  if (!(ty == 1 || ty == 4 || ty == 8) || nt < 0 || nt > 3 || xchunk < 0 || remap < 0 || remap > 2)
    return s->fail(VSE_ERR_ARG, "tuning: ty in {1,4,8}, nt in {0..3}, xchunk >= 0, remap in {0,1,2}");
  if ((nt & 2) && (s->d.ndim != 3 || kernel != 1 || xchunk < 1 || ty < 2 * s->d.ord))
    return s->fail(VSE_ERR_ARG, "tuning: nt bit 1 (LDS-staged y-neighbours) needs the 3D velocity pass (kernel 1), "
                                "xchunk >= 1 and ty >= 2 * FD order");
  if (kernel == 2 && s->d.ndim == 3 && (ty == 1 || nt != 0 || xchunk < 1))
    return s->fail(VSE_ERR_ARG, "fused-step tuning: ty in {4,8}, nt 0, xchunk >= 1");
  if (kernel == 2 && s->d.ndim == 2 && (nt != 0 || !fused2d_fits(s, xchunk)))
    return s->fail(VSE_ERR_ARG, "fused 2D tuning: nt 0, xchunk = x-rows per workgroup, at least ~4 "
                                "(one |v|^2 slot per workgroup)");
  if (s->d.ndim == 2 && kernel != 2 && (xchunk > 16 || nt > 1 || (nt == 1 && kernel != 1)))
    return s->fail(VSE_ERR_ARG, "2D tuning: kernel 0 or 1, xchunk (row groups per workgroup) 0 .. 16, "
                                "nt 1 (x-row march) for the velocity pass only");
  OnDevice on_dev(s->device);
  s->tune[kernel] = KTune{ty, nt, xchunk, remap};
  s->drop_graph();
  return VSE_OK;
}

int vse_set_math(vse_sim *s, int mode) {
  if (check_handle(s) || (mode != VSE_MATH_STRICT && mode != VSE_MATH_FAST)) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  s->math = mode;
  s->drop_graph();
  return VSE_OK;
}

int vse_read_field(vse_sim *s, int field, float *out) {
  if (check_handle(s) || field < 0 || field > 8 || !out) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (!s->field[field]) return s->fail(VSE_ERR_ARG, "field not present in this simulation");
  int rc = s->hip(hipMemcpyAsync(out, s->field[field], s->ncell * sizeof(float), hipMemcpyDeviceToHost, s->stream), "read field");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "read field sync");
  return rc;
}

int vse_read_traces(vse_sim *s, float *out) {
  if (check_handle(s) || !out) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (!s->traces) return s->fail(VSE_ERR_STATE, "no receivers");
  int rc = s->hip(hipMemcpyAsync(out, s->traces, (size_t)s->d.n_rows * s->n_rec * sizeof(float), hipMemcpyDeviceToHost, s->stream), "read traces");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "read traces sync");
  return rc;
}

int vse_read_memory(vse_sim *s, int comp, float *out) {
  if (check_handle(s) || comp < 0 || comp > 5 || !out) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  if (!s->r[comp]) return s->fail(VSE_ERR_ARG, "memory variable not present");
  const int L = s->d.n_sls;
  std::vector<float> planes((size_t)s->rs * L);
  int rc = s->hip(hipMemcpyAsync(planes.data(), s->r[comp], planes.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream), "read r");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "read r sync");
  if (rc) return rc;
  for (long c = 0; c < s->ncell; c++)
    for (int l = 0; l < L; l++) out[c * L + l] = planes[(size_t)l * s->rs + c];
  return 0;
}

int vse_read_v2max(vse_sim *s, float *out) {
  if (check_handle(s) || !out) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  int rc = fold_v2(s);
  if (!rc) rc = s->hip(hipMemcpyAsync(s->host_ctrl, s->ctrl, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream), "read ctrl");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "sync");
  if (!rc) memcpy(out, &s->host_ctrl[0], 4);
  return rc;
}

int vse_get_step(vse_sim *s, int *out) {
  if (check_handle(s) || !out) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  int rc = s->hip(hipMemcpyAsync(s->host_ctrl, s->ctrl, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream), "read ctrl");
  if (!rc) rc = s->hip(hipStreamSynchronize(s->stream), "sync");
  if (!rc) *out = (int)s->host_ctrl[1];
  return rc;
}

void *vse_field_device_ptr(vse_sim *s, int field) {
  if (!s || field < 0 || field > 8) return nullptr;
  return s->field[field];
}

size_t vse_device_bytes(vse_sim *s) { return s ? s->bytes : 0; }

const char *vse_device_name(vse_sim *s) { return s ? s->devname.c_str() : ""; }

const char *vse_last_error(vse_sim *s) { return s ? s->err.c_str() : "null handle"; }

void vse_destroy(vse_sim *s) {
  if (!s) return;
  OnDevice on_dev(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  // detach the batches and gradient handles built on s (they may be destroyed later)
  for (vse_batch *b : s->batches) batch_forget_member(b, s);
  s->batches.clear();
  for (vse_grad *g : s->grads) grad_detach(g);
  s->grads.clear();
  s->drop_graph();
  for (auto &a : s->allocs) (void)hipFree(a.second);
  s->allocs.clear();
  if (s->host_ctrl) (void)hipHostFree(s->host_ctrl);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Many 2D simulations (shots of one specimen, or specimens of different
// sizes) advanced together.  Members are packed, in order, into groups whose
// device state fits a cache budget (the 256 MB Infinity Cache of MI355X by
// default); each group runs all its steps before the next one starts, with
// one stress, one velocity and one sparse launch per step covering every
// member of the group.  Small grids thus fill the chip together, and a
// group's state stays cache-resident from step to step, as a lone
// simulation's does.  A group of one member runs that member's own kernels.
struct BatchGroup {
  int first = 0, n = 0;
  BatchArgs B{};
  int total_tiles = 0;
  bool rec_after = false;     // a member records a velocity channel: one gather launch after each velocity pass
  hipGraphExec_t gexec = nullptr;
  int gexec_steps = 0;
};

struct vse_batch {
  std::vector<vse_sim *> sims;
  int device = 0;
  hipStream_t stream = nullptr;
  int variant = 0, visco = 0, ord = 0;
  int math = VSE_MATH_STRICT;                       // the members' arithmetic mode (they must agree)
  int rows = 1;
  int schedule = VSE_BATCH_SCHED_GRAPH;
  size_t group_bytes = 0;
  std::vector<BatchGroup> groups;
  BatchGroup all;                                   // every member in one launch (vse_batch_profile)
  StepArgs *d_args = nullptr;
  SparseArgs *d_sparse = nullptr;
  unsigned *d_ctrl = nullptr, *h_ctrl = nullptr;   // (v2max bits, step) per member
  unsigned *d_bar = nullptr;                        // grid-barrier words of the persistent schedule
  bool orphan = false;                              // a member was destroyed: the batch can only be destroyed
  std::string err;

  int fail(int code, const std::string &m) { err = m; return code; }
  int hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return VSE_ERR_HIP;
  }
  void drop_graphs() {
    for (auto &g : groups)
      if (g.gexec) { (void)hipGraphExecDestroy(g.gexec); g.gexec = nullptr; g.gexec_steps = 0; }
  }
  int usable() { return orphan ? fail(VSE_ERR_STATE, "a member of this batch was destroyed") : 0; }
};

static void batch_forget_member(vse_batch *b, vse_sim *s) {
  for (auto &m : b->sims)
    if (m == s) m = nullptr;
  b->orphan = true;
}

// Whether each group (and the all-member group) needs the velocity-channel
// gather after its velocity passes; a group whose need changed recaptures its graph.
static void refresh_rec_flags(vse_batch *b) {
  auto need = [&](const BatchGroup &g) {
    for (int i = 0; i < g.n; i++)
      if (velocity_field(b->sims[g.first + i]->rec_channel)) return true;
    return false;
  };
  for (auto &g : b->groups) {
    const bool r = need(g);
    if (r != g.rec_after && g.gexec) { (void)hipGraphExecDestroy(g.gexec); g.gexec = nullptr; g.gexec_steps = 0; }
    g.rec_after = r;
  }
  b->all.rec_after = need(b->all);
}

static int build_group(vse_batch *b, BatchGroup &g, int first, int n) {
  g.first = first;
  g.n = n;
  g.B = BatchArgs{};
  g.B.args = b->d_args + first;
  g.B.sparse = b->d_sparse + first;
  g.B.n = n;
  g.B.L = b->sims[0]->d.n_sls;
  g.B.rows = b->rows;
  long tiles = 0;
  for (int i = 0; i < n; i++) {
    g.B.tile_beg[i] = (int)tiles;
    tiles += vse_tiles2d(b->sims[first + i]->d.nx, b->sims[first + i]->d.ny, b->rows);
    if (tiles > 0x7fffffffL) return b->fail(VSE_ERR_ARG, "batch too large for one launch");
  }
  g.B.tile_beg[n] = (int)tiles;
  g.total_tiles = (int)tiles;
  return 0;
}

// Greedy packing in member order: a group closes when the next member would
// take its device bytes past group_bytes (0: one group of every member).
static int regroup(vse_batch *b) {
  b->drop_graphs();
  b->groups.clear();
  const int n = (int)b->sims.size();
  int rc = build_group(b, b->all, 0, n);
  int first = 0;
  size_t bytes = 0;
  for (int i = 0; i < n && !rc; i++) {
    const size_t mb = b->sims[i]->bytes;
    if (b->group_bytes && i > first && bytes + mb > b->group_bytes) {
      b->groups.emplace_back();
      rc = build_group(b, b->groups.back(), first, i - first);
      first = i;
      bytes = 0;
    }
    bytes += mb;
  }
  if (!rc) {
    b->groups.emplace_back();
    rc = build_group(b, b->groups.back(), first, n - first);
  }
  if (!rc) refresh_rec_flags(b);
  return rc;
}

// Step i of k launched together for every member of g; the velocity pass
// does the members' sources and receivers, and vse_launch_advance_batch
// moves their counters after the k steps.
static hipError_t enqueue_group_step(vse_batch *b, const BatchGroup &g, int i) {
  hipError_t e = vse_launch_batch2d(true, b->variant, b->visco, b->ord, b->math, g.B, g.total_tiles, 0, b->stream);
  if (e == hipSuccess)
    e = vse_launch_batch2d(false, b->variant, b->visco, b->ord, b->math, g.B, g.total_tiles, i, b->stream);
  if (e == hipSuccess && g.rec_after) e = vse_launch_record2d_batch(g.B, i, b->stream);
  return e;
}

// Members' argument blocks as of now (their device buffers may have been
// replaced by vse_set_sources / vse_set_receivers since the last run).
static int upload_batch_args(vse_batch *b) {
  const int n = (int)b->sims.size();
  const int math = b->sims[0]->math;
  for (int i = 1; i < n; i++)
    if (b->sims[i]->math != math) return b->fail(VSE_ERR_ARG, "batch members differ in arithmetic mode");
  if (math != b->math) {
    b->drop_graphs();
    b->math = math;
  }
  std::vector<StepArgs> A(n);
  std::vector<SparseArgs> S(n);
  for (int i = 0; i < n; i++) {
    vse_sim *s = b->sims[i];
    int rc = finalize(s);
    if (rc) return b->fail(rc, "member " + std::to_string(i) + ": " + s->err);
    rc = s->hip(hipStreamSynchronize(s->stream), "member sync");
    if (rc) return b->fail(rc, "member " + std::to_string(i) + ": " + s->err);
    make_args(s, A[i], S[i]);
  }
  SyncOnExit sync(b->stream);     // after the vectors: waits before they are freed
  int rc = b->hip(hipMemcpyAsync(b->d_args, A.data(), n * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream), "upload args");
  if (!rc) rc = b->hip(hipMemcpyAsync(b->d_sparse, S.data(), n * sizeof(SparseArgs), hipMemcpyHostToDevice, b->stream), "upload args");
  if (!rc) rc = b->hip(hipStreamSynchronize(b->stream), "upload args");
  if (!rc) refresh_rec_flags(b);
  return rc;
}

// Stability of the members of g (device max of |v|^2 reached +inf or NaN).
static int check_group(vse_batch *b, const BatchGroup &g) {
  unsigned *dc = b->d_ctrl + 2 * g.first, *hc = b->h_ctrl + 2 * g.first;
  int rc = b->hip(vse_launch_batch_ctrl(g.B, dc, b->stream), "ctrl gather");
  if (!rc) rc = b->hip(hipMemcpyAsync(hc, dc, 2 * g.n * sizeof(unsigned), hipMemcpyDeviceToHost, b->stream), "read ctrl");
  if (!rc) rc = b->hip(hipStreamSynchronize(b->stream), "sync");
  if (rc) return rc;
  for (int i = 0; i < g.n; i++)
    if (hc[2 * i] >= 0x7f800000u) {
      float v;
      memcpy(&v, &hc[2 * i], 4);
      char buf[160];
      snprintf(buf, sizeof buf, "Simulacao tornando-se instavel [member %d: |v|^2 = %g at step %u]", g.first + i, v,
               hc[2 * i + 1]);
      b->sims[g.first + i]->err = buf;
      return b->fail(VSE_ERR_UNSTABLE, buf);
    }
  return 0;
}

// n_steps of one multi-member group on the batch stream, in hipGraphs of up
// to check_every steps.
static int run_group(vse_batch *b, BatchGroup &g, int n_steps, int check_every, float *elapsed_ms) {
  int rc = 0;
  const int chunk = std::max(1, std::min(check_every > 0 ? check_every : 256, std::min(n_steps, 256)));
  if (!(g.gexec && g.gexec_steps == chunk)) {
    if (g.gexec) (void)hipGraphExecDestroy(g.gexec);
    g.gexec = nullptr;
    g.gexec_steps = 0;
    hipGraph_t gr = nullptr;
    rc = b->hip(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    hipError_t e = hipSuccess;
    for (int i = 0; i < chunk && e == hipSuccess; i++) e = enqueue_group_step(b, g, i);
    if (e == hipSuccess) e = vse_launch_advance_batch(g.B, chunk, b->stream);
    hipError_t e2 = hipStreamEndCapture(b->stream, &gr);
    if (e != hipSuccess) { if (gr) (void)hipGraphDestroy(gr); return b->hip(e, "capture launch"); }
    rc = b->hip(e2, "end capture");
    if (!rc) rc = b->hip(hipGraphInstantiate(&g.gexec, gr, nullptr, nullptr, 0), "graph instantiate");
    if (gr) (void)hipGraphDestroy(gr);
    if (rc) return rc;
    g.gexec_steps = chunk;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    rc = b->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = b->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = b->hip(hipEventRecord(e0, b->stream), "event record");
  }
  int done = 0, since_check = 0;
  while (!rc && done < n_steps) {
    const int left = n_steps - done;
    if (left >= chunk) {
      rc = b->hip(hipGraphLaunch(g.gexec, b->stream), "graph launch");
      done += chunk;
      since_check += chunk;
    } else {
      for (int i = 0; i < left && !rc; i++) rc = b->hip(enqueue_group_step(b, g, i), "step launch");
      if (!rc) rc = b->hip(vse_launch_advance_batch(g.B, left, b->stream), "step launch");
      done += left;
      since_check += left;
    }
    if (!rc && check_every > 0 && since_check >= check_every && done < n_steps) {
      rc = check_group(b, g);
      since_check = 0;
    }
  }
  if (elapsed_ms && !rc) {
    rc = b->hip(hipEventRecord(e1, b->stream), "event record");
    if (!rc) rc = b->hip(hipEventSynchronize(e1), "event sync");
    if (!rc) rc = b->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = check_group(b, g);
  else (void)hipStreamSynchronize(b->stream);
  return rc;
}

// Steps per cooperative launch of the persistent schedule at most (a launch
// is one kernel; this bounds how long one runs).
constexpr int kPersistentMaxSteps = 1024;

// n_steps of group g (one member or many) under the persistent schedule:
// cooperative launches of up to check_every steps, each running every step
// of every member of the group with grid-wide barriers between the passes.
static int run_group_persistent(vse_batch *b, BatchGroup &g, int n_steps, int check_every, float *elapsed_ms) {
  int grid = 0;
  int rc = b->hip(vse_batch2d_persistent_grid(b->variant, b->visco, b->ord, b->math, g.B.L, g.total_tiles, &grid),
                  "persistent launch shape (cooperative launch)");
  if (rc) return rc;
  const int chunk = std::min(kPersistentMaxSteps, check_every > 0 ? check_every : kPersistentMaxSteps);
  unsigned *const fault = b->d_bar + VSE_GRID_BAR_FAULT;
  unsigned *const h_fault = b->h_ctrl + 2 * b->sims.size();
  // a barrier that gave up leaves the fault word set: read it back after the launches
  auto faulted = [&]() -> int {
    int r = b->hip(hipMemcpyAsync(h_fault, fault, sizeof(unsigned), hipMemcpyDeviceToHost, b->stream), "read fault");
    if (!r) r = b->hip(hipStreamSynchronize(b->stream), "sync");
    if (!r && *h_fault) r = b->fail(VSE_ERR_HIP, "persistent launch: a grid barrier wait exceeded its limit");
    return r;
  };
  rc = b->hip(hipMemsetAsync(fault, 0, sizeof(unsigned), b->stream), "clear fault");
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    rc = b->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = b->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = b->hip(hipEventRecord(e0, b->stream), "event record");
  }
  int done = 0, since_check = 0;
  while (!rc && done < n_steps) {
    const int k = std::min(chunk, n_steps - done);
    rc = b->hip(vse_launch_batch2d_persistent(b->variant, b->visco, b->ord, b->math, g.B, g.total_tiles, grid, k,
                                              g.rec_after, b->d_bar, b->stream),
                "persistent launch");
    done += k;
    since_check += k;
    if (!rc && check_every > 0 && since_check >= check_every && done < n_steps) {
      rc = faulted();
      if (!rc) rc = check_group(b, g);
      since_check = 0;
    }
  }
  if (elapsed_ms && !rc) {
    rc = b->hip(hipEventRecord(e1, b->stream), "event record");
    if (!rc) rc = b->hip(hipEventSynchronize(e1), "event sync");
    if (!rc) rc = b->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = faulted();
  if (!rc) rc = check_group(b, g);
  else (void)hipStreamSynchronize(b->stream);
  return rc;
}

extern "C" {

void vse_batch_destroy(vse_batch *b) {
  if (!b) return;
  OnDevice on_dev(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (vse_sim *s : b->sims)
    if (s) s->batches.erase(std::remove(s->batches.begin(), s->batches.end(), b), s->batches.end());
  b->drop_graphs();
  if (b->d_args) (void)hipFree(b->d_args);
  if (b->d_sparse) (void)hipFree(b->d_sparse);
  if (b->d_ctrl) (void)hipFree(b->d_ctrl);
  if (b->d_bar) (void)hipFree(b->d_bar);
  if (b->h_ctrl) (void)hipHostFree(b->h_ctrl);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int vse_batch_create(vse_sim *const *sims, int n, vse_batch **out) {
  static_assert(VSE_MAX_BATCH == VSE_BATCH_MAX, "batch limit");
  if (!sims || !out || n < 1 || n > VSE_BATCH_MAX) return VSE_ERR_ARG;
  *out = nullptr;
  vse_sim *s0 = sims[0];
  if (!s0) return VSE_ERR_ARG;
  OnDevice on_dev(s0->device);
  for (int i = 0; i < n; i++) {
    vse_sim *s = sims[i];
    if (!s || s->d.ndim != 2 || s->d.variant != s0->d.variant || s->d.visco != s0->d.visco ||
        s->d.ord != s0->d.ord || s->d.n_sls != s0->d.n_sls || s->device != s0->device)
      return VSE_ERR_ARG;   // members: 2D, one device, one kernel instantiation
    for (int j = 0; j < i; j++)
      if (sims[j] == s) return VSE_ERR_ARG;
  }
  vse_batch *b = new vse_batch();
  b->sims.assign(sims, sims + n);
  for (vse_sim *s : b->sims) s->batches.push_back(b);   // vse_batch_destroy (also on failure below) unregisters
  b->device = s0->device;
  b->variant = s0->d.variant;
  b->visco = s0->d.visco;
  b->ord = s0->d.ord;
  b->group_bytes = VSE_BATCH_GROUP_BYTES_DEFAULT;
  int rc = b->hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!rc) rc = b->hip(hipMalloc((void **)&b->d_args, n * sizeof(StepArgs)), "hipMalloc");
  if (!rc) rc = b->hip(hipMalloc((void **)&b->d_sparse, n * sizeof(SparseArgs)), "hipMalloc");
  if (!rc) rc = b->hip(hipMalloc((void **)&b->d_ctrl, 2 * n * sizeof(unsigned)), "hipMalloc");
  if (!rc) rc = b->hip(hipHostMalloc((void **)&b->h_ctrl, (2 * n + 1) * sizeof(unsigned), hipHostMallocDefault),
                       "hipHostMalloc");
  if (!rc) rc = b->hip(hipMalloc((void **)&b->d_bar, VSE_GRID_BAR_WORDS * sizeof(unsigned)), "hipMalloc");
  if (!rc) rc = upload_batch_args(b);   // finalizes the members: their device bytes are complete
  if (!rc) rc = regroup(b);
  if (rc) {
    vse_batch_destroy(b);
    return rc;
  }
  *out = b;
  return VSE_OK;
}

int vse_batch_run(vse_batch *b, int n_steps, int check_every, float *elapsed_ms) {
  if (!b || n_steps < 0) return VSE_ERR_ARG;
  OnDevice on_dev(b->device);
  int rc = b->usable();
  if (!rc) rc = upload_batch_args(b);
  if (rc) return rc;
  double total = 0.0;
  for (size_t k = 0; k < b->groups.size() && !rc && n_steps > 0; k++) {
    BatchGroup &g = b->groups[k];
    float ms = 0.f;
    if (b->schedule == VSE_BATCH_SCHED_PERSISTENT) {
      rc = run_group_persistent(b, g, n_steps, check_every, elapsed_ms ? &ms : nullptr);
    } else if (g.n == 1) {
      vse_sim *s = b->sims[g.first];
      rc = vse_run(s, n_steps, check_every, elapsed_ms ? &ms : nullptr);
      if (rc) {
        b->err = "member " + std::to_string(g.first) + ": " + s->err;
        if (rc == VSE_ERR_UNSTABLE)
          b->err = "Simulacao tornando-se instavel [member " + std::to_string(g.first) + ": " + s->err + "]";
      }
    } else {
      rc = run_group(b, g, n_steps, check_every, elapsed_ms ? &ms : nullptr);
    }
    total += ms;
  }
  if (elapsed_ms) *elapsed_ms = rc ? 0.f : (float)total;
  return rc;
}

int vse_batch_profile(vse_batch *b, int n_steps, float *ms) {
  if (!b || n_steps < 1 || !ms) return VSE_ERR_ARG;
  OnDevice on_dev(b->device);
  int rc = b->usable();
  if (!rc) rc = upload_batch_args(b);
  if (rc) return rc;
  const BatchGroup &g = b->all;
  hipEvent_t ev[4] = {};
  for (int i = 0; i < 4 && !rc; i++) rc = b->hip(hipEventCreate(&ev[i]), "event");
  double acc[3] = {0, 0, 0};
  for (int it = 0; it < n_steps && !rc; it++) {
    hipError_t e = hipEventRecord(ev[0], b->stream);
    if (e == hipSuccess)
      e = vse_launch_batch2d(true, b->variant, b->visco, b->ord, b->math, g.B, g.total_tiles, 0, b->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
    if (e == hipSuccess)
      e = vse_launch_batch2d(false, b->variant, b->visco, b->ord, b->math, g.B, g.total_tiles, 0, b->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[2], b->stream);
    if (e == hipSuccess && g.rec_after) e = vse_launch_record2d_batch(g.B, 0, b->stream);
    if (e == hipSuccess) e = vse_launch_advance_batch(g.B, 1, b->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[3], b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[3]);
    rc = b->hip(e, "profile step");
    for (int k = 0; k < 3 && !rc; k++) {
      float t = 0.f;
      rc = b->hip(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "elapsed");
      acc[k] += t;
    }
  }
  for (int i = 0; i < 4; i++)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (!rc)
    for (int k = 0; k < 3; k++) ms[k] = (float)(acc[k] / n_steps);
  if (!rc) rc = check_group(b, g);
  return rc;
}

int vse_batch_set_rows(vse_batch *b, int rows) {
  if (!b || rows < 1 || rows > 64) return VSE_ERR_ARG;
  OnDevice on_dev(b->device);
  if (int rc = b->usable()) return rc;
  b->rows = rows;
  return regroup(b);
}

int vse_batch_set_group_bytes(vse_batch *b, size_t bytes) {
  if (!b) return VSE_ERR_ARG;
  OnDevice on_dev(b->device);
  if (int rc = b->usable()) return rc;
  b->group_bytes = bytes;
  return regroup(b);
}

int vse_batch_set_schedule(vse_batch *b, int schedule) {
  if (!b || (schedule != VSE_BATCH_SCHED_GRAPH && schedule != VSE_BATCH_SCHED_PERSISTENT)) return VSE_ERR_ARG;
  OnDevice on_dev(b->device);
  if (int rc = b->usable()) return rc;
  if (schedule == VSE_BATCH_SCHED_PERSISTENT) {
    int coop = 0;
    int rc = b->hip(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b->device), "device attribute");
    if (rc) return rc;
    if (!coop) return b->fail(VSE_ERR_HIP, "the device cannot launch cooperative kernels");
  }
  b->schedule = schedule;
  return VSE_OK;
}

int vse_batch_groups(vse_batch *b, int *first, int max_groups) {
  if (!b) return 0;
  const int ng = (int)b->groups.size();
  if (first)
    for (int k = 0; k < ng && k < max_groups; k++) first[k] = b->groups[k].first;
  return ng;
}

int vse_batch_size(vse_batch *b) { return b ? (int)b->sims.size() : 0; }

const char *vse_batch_last_error(vse_batch *b) { return b ? b->err.c_str() : "null handle"; }

}  // extern "C"

// ------------------------------------------------------- gradient (2D visco)
// Checkpointed adjoint of one 2D viscoelastic simulation (WGSL variant).  The
// forward pass keeps the full state every K steps (K = ceil(sqrt(n_rows)) by
// default); the reverse sweep recomputes each K-step segment from its
// checkpoint into K stored levels, then runs the adjoint steps of that
// segment backwards.  Device memory: (ceil(n_rows/K) + K) levels of
// (5 + 3L) fields plus the CPML bands, one adjoint level, 8 derivative
// adjoints and 6 FP64 gradient maps.
struct vse_grad {
  vse_sim *s = nullptr;
  int K = 0, n_ckpt = 0, n_fwd = 0;
  int flush = VSE_GRAD_FLUSH;     // steps per FP64 flush of the FP32 partials
  int n_rec = 0;                  // receivers of s when the trace buffers were sized
  bool last_seg_stored = false;   // the forward left the last segment's levels in seg[]
  std::vector<DLevel> ckpt, seg;
  // With sqrt(N) checkpoints the reverse sweep recomputes segment sg - 1 on a
  // second stream (rstream) while the adjoint of segment sg runs on the
  // simulation's stream; consecutive segments alternate between seg and seg2.
  std::vector<DLevel> seg2;
  int top = 0;                    // last segment of the latest forward (its levels are in seg)
  bool overlap = true;
  hipStream_t rstream = nullptr;
  hipEvent_t ev_go = nullptr, ev_rec[2] = {}, ev_adj[2] = {};
  DLevel bar{};
  float *abar[8] = {};
  double *g[VSE_GR_COUNT] = {};
  float *gf[VSE_GR_COUNT] = {};   // FP32 per-step partials of g
  float *ts = nullptr, *traces = nullptr, *gbar = nullptr, *obs = nullptr;
  bool ts_unit = true;            // every tau scale is exactly 1
  int *src_map = nullptr;
  double *misfit = nullptr;
  unsigned *ctrl = nullptr, *host_ctrl = nullptr;
  long rs = 0;                    // mechanism stride of the memory-variable planes (floats)
  std::vector<void *> allocs;
  size_t bytes = 0;
  std::string err;

  int fail(int code, const std::string &m) { err = m; return code; }
  int hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return VSE_ERR_HIP;
  }
  // Zeroed device array of n elements, its base skewed by 256 B modulo 4 KiB
  // from the previous array of the process, as the simulation's arrays are:
  // the differentiable kernels stream 20-50 arrays at the same cell index,
  // and with every base on one 4 KiB alignment those streams meet in the same
  // HBM channels.
  template <typename T>
  int alloc(T **p, size_t n) {
    constexpr size_t kSkewStep = 256, kSkewSpan = 16;
    const size_t nb = std::max<size_t>(n, 1) * sizeof(T);
    const size_t skew = kSkewStep * (vse_sim::g_allocs.fetch_add(1, std::memory_order_relaxed) % kSkewSpan);
    void *q = nullptr;
    int rc = hip(hipMalloc(&q, nb + kSkewStep * kSkewSpan), "hipMalloc");
    if (rc) return rc;
    allocs.push_back(q);
    bytes += nb;
    *p = (T *)((char *)q + skew);
    return hip(hipMemsetAsync(*p, 0, nb, s->stream), "hipMemsetAsync");
  }
  long psi_elems(int i) const {   // local slot i of kPsiSlots2D: even = x axis, odd = y axis
    return (i % 2 == 0) ? (long)s->ax[0].nb * s->d.ny : (long)s->d.nx * s->ax[1].nb;
  }
  // A level is one allocation.  Its arrays follow each other with a 256 B gap,
  // and the mechanism planes of a memory variable are rs floats apart (1.5 KiB
  // more than a plane), so no two streams of a cell start on the same 4 KiB
  // alignment (the simulation's own layout does the same).
  static constexpr long kGap = 64;
  long level_floats() const {
    long total = 5 * (s->ncell + kGap) + 3 * ((long)s->d.n_sls * rs + kGap);
    for (int i = 0; i < 8; i++) total += psi_elems(i) + kGap;
    return total;
  }
  int alloc_level(DLevel &v) {
    const long nc = s->ncell;
    const int L = s->d.n_sls;
    float *base = nullptr;
    int rc = alloc(&base, (size_t)level_floats());
    if (rc) return rc;
    float *p = base;
    float **f[5] = {&v.vx, &v.vy, &v.sxx, &v.syy, &v.sxy};
    for (int i = 0; i < 5; i++) { *f[i] = p; p += nc + kGap; }
    for (int c = 0; c < 3; c++) { v.r[c] = p; p += (long)L * rs + kGap; }
    for (int i = 0; i < 8; i++) { v.psi[i] = p; p += psi_elems(i) + kGap; }
    return 0;
  }
  int zero_level(const DLevel &v) {
    return hip(hipMemsetAsync(v.vx, 0, (size_t)level_floats() * sizeof(float), s->stream), "zero level");
  }
  // level j of segment sg (state sg*K + j)
  const DLevel &lev(int sg, int j) const {
    if (j == 0) return ckpt[sg];
    return (!seg2.empty() && ((top - sg) & 1)) ? seg2[j - 1] : seg[j - 1];
  }
};

static void make_diff(vse_grad *g, DiffArgs &A) {
  vse_sim *s = g->s;
  const vse_desc &d = s->d;
  memset(&A, 0, sizeof(A));
  A.nx = d.nx; A.ny = d.ny; A.ord = d.ord; A.ncell = s->ncell; A.rs = g->rs;
  A.dx = d.dx; A.dy = d.dy; A.dt = d.dt;
  for (int c = 0; c < d.ord; c++) {
    A.C[c] = d.coefs[c];
    A.Cd[0][c] = d.coefs[c] / d.dx;
    A.Cd[1][c] = d.coefs[c] / d.dy;
  }
  A.ax[0] = s->ax[0]; A.ax[1] = s->ax[1];
  A.rho = s->rho; A.cp = s->cp; A.cs = s->cs; A.ts = g->ts; A.ts_unit = g->ts_unit ? 1 : 0;
  A.L = d.n_sls;
  A.zk = s->zk;
  A.bar = g->bar;
  for (int i = 0; i < 8; i++) A.abar[i] = g->abar[i];
  for (int i = 0; i < VSE_GR_COUNT; i++) {
    A.g[i] = g->g[i];
    A.gf[i] = g->gf[i];
  }
  A.v2max = g->ctrl;
  A.src_map = g->src_map; A.src = s->src; A.src_cols = s->src_cols; A.src_rows = d.n_rows;
  A.src_rowflag = s->src_rowflag;
  A.n_rec = s->n_rec; A.rec_gate = s->rec_gate; A.rec_beg = s->rec_beg; A.rec_cell = s->rec_cell;
  A.rec_delay = s->rec_delay; A.traces = g->traces; A.gbar = g->gbar;
}

// One out-of-place forward step n: level `in` -> level `out` (two launches
// on `st`, the simulation's stream by default).
static int grad_step(vse_grad *g, DiffArgs &A, const DLevel &in, const DLevel &out, int n, bool record,
                     hipStream_t st = nullptr) {
  A.in = in; A.out = out;
  A.n = n; A.record = record ? 1 : 0;
  if (!st) st = g->s->stream;
  hipError_t e = vse_launch_dstress(A, st);
  if (e == hipSuccess) e = vse_launch_dvelocity(A, st);
  return g->hip(e, "forward step");
}

static int grad_check(vse_grad *g, int n) {
  int rc = g->hip(hipMemcpyAsync(g->host_ctrl, g->ctrl, sizeof(unsigned), hipMemcpyDeviceToHost, g->s->stream), "ctrl");
  if (!rc) rc = g->hip(hipStreamSynchronize(g->s->stream), "sync");
  if (!rc && g->host_ctrl[0] >= 0x7f800000u) {
    char buf[128];
    snprintf(buf, sizeof buf, "Simulacao tornando-se instavel [gradient forward, before step %d]", n);
    return g->fail(VSE_ERR_UNSTABLE, buf);
  }
  return rc;
}

// The per-cell source-column map from the simulation's current source list.
static int grad_src_map(vse_grad *g) {
  vse_sim *s = g->s;
  return g->hip(vse_launch_src_map(g->src_map, s->ncell, s->src_cell, s->src_col, s->n_src_pts, s->stream), "source map");
}

extern "C" {

// Releases g's device memory (on its simulation's device and stream) and
// leaves it without a simulation: every later call but vse_grad_destroy
// returns VSE_ERR_STATE.
static void grad_detach(vse_grad *g) {
  vse_sim *s = g->s;
  if (!s) return;
  OnDevice on_dev(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (g->rstream) {
    (void)hipStreamSynchronize(g->rstream);
    (void)hipStreamDestroy(g->rstream);
    g->rstream = nullptr;
  }
  hipEvent_t *evs[5] = {&g->ev_go, &g->ev_rec[0], &g->ev_rec[1], &g->ev_adj[0], &g->ev_adj[1]};
  for (hipEvent_t *e : evs)
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  for (void *p : g->allocs) (void)hipFree(p);
  g->allocs.clear();
  if (g->host_ctrl) (void)hipHostFree(g->host_ctrl);
  g->host_ctrl = nullptr;
  g->s = nullptr;
}

// 0 if g can run, else VSE_ERR_ARG (null handle) or VSE_ERR_STATE (its
// simulation was destroyed, or its receivers changed after the trace buffers
// were sized)
static int grad_usable(vse_grad *g) {
  if (!g) return VSE_ERR_ARG;
  if (!g->s) return g->fail(VSE_ERR_STATE, "the simulation of this gradient handle was destroyed");
  if (g->s->n_rec != g->n_rec || g->s->rec_channel != VSE_F_SYY)
    return g->fail(VSE_ERR_STATE, "the simulation's receivers changed after vse_grad_create");
  return 0;
}

void vse_grad_destroy(vse_grad *g) {
  if (!g) return;
  if (vse_sim *s = g->s) s->grads.erase(std::remove(s->grads.begin(), s->grads.end(), g), s->grads.end());
  grad_detach(g);
  delete g;
}

int vse_grad_create(vse_sim *s, int seg_len, vse_grad **out) {
  if (!s || !out || seg_len < 0 || seg_len == 1) return VSE_ERR_ARG;
  OnDevice on_dev(s->device);
  *out = nullptr;
  if (s->d.ndim != 2 || s->d.variant != VSE_VARIANT_WGSL || !s->d.visco)
    return s->fail(VSE_ERR_ARG, "gradient: 2D viscoelastic WGSL-variant simulations only");
  if (s->rec_channel != VSE_F_SYY)
    return s->fail(VSE_ERR_ARG, "gradient: receivers must record sigma_yy (the adjoint injects its residual there)");
  int rc = finalize(s);
  if (rc) return rc;
  vse_grad *g = new vse_grad();
  g->s = s;
  g->rs = s->ncell + 384;
  s->grads.push_back(g);      // vse_grad_destroy (also on failure below) unregisters
  const int N = s->d.n_rows;
  g->K = seg_len ? seg_len : std::max(2, (int)std::ceil(std::sqrt((double)N)));
  g->n_ckpt = (N + g->K - 1) / g->K;
  g->ckpt.resize(g->n_ckpt);
  g->seg.resize(g->K);
  for (auto &v : g->ckpt)
    if (!rc) rc = g->alloc_level(v);
  for (auto &v : g->seg)
    if (!rc) rc = g->alloc_level(v);
  if (g->n_ckpt > 1) {     // the second segment buffer, stream and events of the overlapped reverse sweep
    g->seg2.resize(g->K);
    for (auto &v : g->seg2)
      if (!rc) rc = g->alloc_level(v);
    if (!rc) rc = g->hip(hipStreamCreateWithFlags(&g->rstream, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t *evs[5] = {&g->ev_go, &g->ev_rec[0], &g->ev_rec[1], &g->ev_adj[0], &g->ev_adj[1]};
    for (hipEvent_t *e : evs)
      if (!rc) rc = g->hip(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
  }
  if (!rc) rc = g->alloc_level(g->bar);
  for (int i = 0; i < 8 && !rc; i++) rc = g->alloc(&g->abar[i], s->ncell);
  for (int i = 0; i < VSE_GR_COUNT && !rc; i++) rc = g->alloc(&g->g[i], s->ncell);
  for (int i = 0; i < VSE_GR_COUNT && !rc; i++) rc = g->alloc(&g->gf[i], s->ncell);
  if (!rc) rc = g->alloc(&g->ts, s->ncell);
  if (!rc) rc = g->alloc(&g->src_map, s->ncell);
  g->n_rec = s->n_rec;
  const size_t nt = (size_t)N * std::max(s->n_rec, 1);
  if (!rc) rc = g->alloc(&g->traces, nt);
  if (!rc) rc = g->alloc(&g->gbar, nt);
  if (!rc) rc = g->alloc(&g->obs, nt);
  if (!rc) rc = g->alloc(&g->misfit, 1);
  if (!rc) rc = g->alloc(&g->ctrl, 1);
  if (!rc) rc = g->hip(hipHostMalloc((void **)&g->host_ctrl, sizeof(unsigned), hipHostMallocDefault), "hipHostMalloc");
  if (!rc) {
    std::vector<float> ones(s->ncell, 1.0f);
    rc = g->hip(hipMemcpyAsync(g->ts, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice, s->stream), "tau scale");
    if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "create");
  }
  if (rc) {
    s->err = g->err;
    vse_grad_destroy(g);
    return rc;
  }
  *out = g;
  return VSE_OK;
}

int vse_grad_set_tau_scale(vse_grad *g, const float *scale) {
  if (int rc = grad_usable(g)) return rc;
  OnDevice on_dev(g->s->device);
  vse_sim *s = g->s;
  std::vector<float> ones;
  if (!scale) {
    ones.assign(s->ncell, 1.0f);
    scale = ones.data();
  }
  bool unit = true;
  for (long i = 0; i < s->ncell; i++) {
    if (!(scale[i] > 0.f) || !std::isfinite(scale[i])) return g->fail(VSE_ERR_ARG, "tau scale must be positive and finite");
    unit = unit && scale[i] == 1.0f;
  }
  int rc = g->hip(hipMemcpyAsync(g->ts, scale, s->ncell * sizeof(float), hipMemcpyHostToDevice, s->stream), "tau scale");
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "tau scale");
  if (!rc) g->ts_unit = unit;
  return rc;
}

int vse_grad_forward(vse_grad *g, int n_steps, int check_every, float *traces, float *elapsed_ms) {
  if (int rc = grad_usable(g)) return rc;
  if (n_steps < 1 || n_steps > g->s->d.n_rows) return VSE_ERR_ARG;
  OnDevice on_dev(g->s->device);
  vse_sim *s = g->s;
  int rc = finalize(s);
  if (rc) return g->fail(rc, s->err);
  DiffArgs A;
  make_diff(g, A);
  rc = g->zero_level(g->ckpt[0]);
  const size_t nt = (size_t)s->d.n_rows * std::max(s->n_rec, 1);
  if (!rc) rc = g->hip(hipMemsetAsync(g->traces, 0, nt * sizeof(float), s->stream), "zero traces");
  if (!rc) rc = g->hip(hipMemsetAsync(g->ctrl, 0, sizeof(unsigned), s->stream), "zero ctrl");
  if (!rc) rc = grad_src_map(g);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && elapsed_ms) {
    rc = g->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = g->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = g->hip(hipEventRecord(e0, s->stream), "event");
  }
  // Steps before the last segment run in place on one working level (the
  // kernels read a cell's own values before writing them, so in == out is
  // valid, and in place keeps half as many address streams open), except
  // that a step reaching a multiple of K writes its checkpoint out of place.
  // The last segment's steps write straight into its segment levels, which
  // the reverse sweep then uses without recomputing them.
  const int last = (n_steps - 1) / g->K;      // index of the last segment
  g->top = last;                              // its levels go to seg (lev())
  const DLevel *cur = &g->ckpt[0];
  for (int n = 0; n < n_steps && !rc; n++) {
    const int m = n + 1;
    const DLevel *nxt;
    if (n >= last * g->K) nxt = &g->lev(last, m - last * g->K);
    else if (m % g->K == 0) nxt = &g->ckpt[m / g->K];
    else nxt = &g->seg[0];
    rc = grad_step(g, A, *cur, *nxt, n, true);
    cur = nxt;
    if (!rc && check_every > 0 && m % check_every == 0 && m < n_steps) rc = grad_check(g, m);
  }
  if (!rc && elapsed_ms) {
    rc = g->hip(hipEventRecord(e1, s->stream), "event");
    if (!rc) rc = g->hip(hipEventSynchronize(e1), "event");
    if (!rc) rc = g->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = grad_check(g, n_steps);
  if (!rc && traces)
    rc = g->hip(hipMemcpyAsync(traces, g->traces, (size_t)n_steps * s->n_rec * sizeof(float), hipMemcpyDeviceToHost,
                               s->stream), "read traces");
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "forward");
  g->n_fwd = rc ? 0 : n_steps;
  g->last_seg_stored = !rc;
  return rc;
}

int vse_grad_misfit(vse_grad *g, const float *observed, double *misfit) {
  if (int rc = grad_usable(g)) return rc;
  if (!observed || !misfit) return VSE_ERR_ARG;
  OnDevice on_dev(g->s->device);
  if (g->n_fwd < 1) return g->fail(VSE_ERR_STATE, "vse_grad_forward not run");
  vse_sim *s = g->s;
  const long n = (long)g->n_fwd * s->n_rec;
  int rc = g->hip(hipMemcpyAsync(g->obs, observed, n * sizeof(float), hipMemcpyHostToDevice, s->stream), "observed");
  if (!rc) rc = g->hip(hipMemsetAsync(g->misfit, 0, sizeof(double), s->stream), "zero misfit");
  if (!rc && n > 0) rc = g->hip(vse_launch_misfit(g->traces, g->obs, g->gbar, n, g->misfit, s->stream), "misfit");
  if (!rc) rc = g->hip(hipMemcpyAsync(misfit, g->misfit, sizeof(double), hipMemcpyDeviceToHost, s->stream), "misfit");
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "misfit");
  return rc;
}

int vse_grad_set_overlap(vse_grad *g, int overlap) {
  if (int rc = grad_usable(g)) return rc;
  if (overlap != 0 && overlap != 1) return VSE_ERR_ARG;
  g->overlap = overlap != 0;
  return VSE_OK;
}

int vse_grad_backward(vse_grad *g, const float *grad_traces, float *elapsed_ms) {
  if (int rc = grad_usable(g)) return rc;
  OnDevice on_dev(g->s->device);
  const int N = g->n_fwd;
  if (N < 1) return g->fail(VSE_ERR_STATE, "vse_grad_forward not run");
  vse_sim *s = g->s;
  // the velocity adjoint reads the source row table (a failed vse_set_sources
  // since the forward leaves none)
  if (!s->src || !s->src_rowflag) return g->fail(VSE_ERR_STATE, "vse_set_sources not called");
  int rc = 0;
  if (grad_traces)
    rc = g->hip(hipMemcpyAsync(g->gbar, grad_traces, (size_t)N * s->n_rec * sizeof(float), hipMemcpyHostToDevice,
                               s->stream), "adjoint source");
  if (!rc) rc = g->zero_level(g->bar);
  for (int i = 0; i < 8 && !rc; i++)
    rc = g->hip(hipMemsetAsync(g->abar[i], 0, s->ncell * sizeof(float), s->stream), "zero abar");
  for (int i = 0; i < VSE_GR_COUNT && !rc; i++) {
    rc = g->hip(hipMemsetAsync(g->g[i], 0, s->ncell * sizeof(double), s->stream), "zero gradient");
    if (!rc) rc = g->hip(hipMemsetAsync(g->gf[i], 0, s->ncell * sizeof(float), s->stream), "zero partials");
  }
  if (!rc) rc = grad_src_map(g);
  DiffArgs A;
  make_diff(g, A);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && elapsed_ms) {
    rc = g->hip(hipEventCreate(&e0), "event");
    if (!rc) rc = g->hip(hipEventCreate(&e1), "event");
    if (!rc) rc = g->hip(hipEventRecord(e0, s->stream), "event");
  }
  const int nseg = (N + g->K - 1) / g->K;
  // Overlapped sweep: segment sg - 1 is recomputed on the second stream while
  // the adjoint of segment sg runs on the simulation's stream.  Segment sg's
  // levels are in seg or seg2 by the parity of top - sg; a recompute waits for
  // the adjoint that last read its buffer (segment sg + 1, event ev_adj), and
  // an adjoint waits for its segment's recompute (event ev_rec).  The same
  // kernels on the same inputs: the gradient is the same either way.
  const bool ov = g->overlap && g->rstream && nseg > 1;
  DiffArgs Ar = A;                             // the recompute's arguments
  auto par = [&](int sg) { return (g->top - sg) & 1; };
  auto recompute = [&](int sg, hipStream_t st) {
    const int n0 = sg * g->K, n1 = std::min(n0 + g->K, N);
    int r = 0;
    for (int j = 0; j < n1 - n0 && !r; j++) r = grad_step(g, Ar, g->lev(sg, j), g->lev(sg, j + 1), n0 + j, false, st);
    return r;
  };
  if (!rc && ov) {     // the second stream starts after everything issued so far on the simulation's stream
    rc = g->hip(hipEventRecord(g->ev_go, s->stream), "event");
    if (!rc) rc = g->hip(hipStreamWaitEvent(g->rstream, g->ev_go, 0), "stream wait");
  }
  for (int sg = nseg - 1; sg >= 0 && !rc; sg--) {
    const int n0 = sg * g->K, n1 = std::min(n0 + g->K, N);
    const bool stored = sg == nseg - 1 && g->last_seg_stored;
    if (!ov) {
      if (!stored) rc = recompute(sg, s->stream);
    } else {
      if (sg == nseg - 1 && !stored) {
        rc = recompute(sg, g->rstream);
        if (!rc) rc = g->hip(hipEventRecord(g->ev_rec[par(sg)], g->rstream), "event");
      }
      if (!rc && sg >= 1) {
        if (sg + 1 <= nseg - 1) rc = g->hip(hipStreamWaitEvent(g->rstream, g->ev_adj[par(sg + 1)], 0), "stream wait");
        if (!rc) rc = recompute(sg - 1, g->rstream);
        if (!rc) rc = g->hip(hipEventRecord(g->ev_rec[par(sg - 1)], g->rstream), "event");
      }
      if (!rc && !stored) rc = g->hip(hipStreamWaitEvent(s->stream, g->ev_rec[par(sg)], 0), "stream wait");
    }
    for (int n = n1 - 1; n >= n0 && !rc; n--) {
      A.in = g->lev(sg, n - n0);
      A.out = g->lev(sg, n - n0 + 1);
      A.n = n;
      hipError_t e = vse_launch_adj_velocity(A, s->stream);
      if (e == hipSuccess) e = vse_launch_adj_stress(A, s->stream);
      if (e == hipSuccess && n % g->flush == 0) e = vse_launch_grad_flush(A, s->stream);
      rc = g->hip(e, "adjoint step");
    }
    if (!rc && ov) rc = g->hip(hipEventRecord(g->ev_adj[par(sg)], s->stream), "event");
  }
  if (g->rstream) (void)hipStreamSynchronize(g->rstream);    // nothing queued there outlives this call
  g->last_seg_stored = false;    // the recomputed segments overwrote it
  if (!rc && elapsed_ms) {
    rc = g->hip(hipEventRecord(e1, s->stream), "event");
    if (!rc) rc = g->hip(hipEventSynchronize(e1), "event");
    if (!rc) rc = g->hip(hipEventElapsedTime(elapsed_ms, e0, e1), "elapsed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "backward");
  return rc;
}

int vse_grad_read(vse_grad *g, int which, double *out) {
  if (int rc = grad_usable(g)) return rc;
  if (which < 0 || which >= VSE_GR_COUNT || !out) return VSE_ERR_ARG;
  OnDevice on_dev(g->s->device);
  vse_sim *s = g->s;
  int rc = g->hip(hipMemcpyAsync(out, g->g[which], s->ncell * sizeof(double), hipMemcpyDeviceToHost, s->stream), "read gradient");
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "read gradient");
  return rc;
}

int vse_grad_read_state(vse_grad *g, int field, float *out) {
  if (int rc = grad_usable(g)) return rc;
  if (field < 0 || field > 4 || !out) return VSE_ERR_ARG;
  OnDevice on_dev(g->s->device);
  if (g->n_fwd < 1 || !g->last_seg_stored) return g->fail(VSE_ERR_STATE, "no forward state (run vse_grad_forward)");
  const int m = g->n_fwd, last = (m - 1) / g->K;
  const DLevel &lv = g->lev(last, m - last * g->K);    // where the forward ended
  const float *src[5] = {lv.vx, lv.vy, lv.sxx, lv.syy, lv.sxy};
  vse_sim *s = g->s;
  int rc = g->hip(hipMemcpyAsync(out, src[field], s->ncell * sizeof(float), hipMemcpyDeviceToHost, s->stream), "read state");
  if (!rc) rc = g->hip(hipStreamSynchronize(s->stream), "read state");
  return rc;
}

int vse_grad_set_flush(vse_grad *g, int flush_every) {
  if (int rc = grad_usable(g)) return rc;
  if (flush_every < 1 || flush_every > VSE_GRAD_FLUSH_MAX)
    return g->fail(VSE_ERR_ARG, "flush interval must be 1 .. " + std::to_string(VSE_GRAD_FLUSH_MAX));
  g->flush = flush_every;
  return 0;
}

int vse_grad_segment(vse_grad *g) { return g ? g->K : 0; }

size_t vse_grad_device_bytes(vse_grad *g) { return g ? g->bytes : 0; }

const char *vse_grad_last_error(vse_grad *g) { return g ? g->err.c_str() : "null handle"; }

}  // extern "C"

// ebiv_runtime.cpp — host runtime of the MI355X EBIV decoder behind the C ABI
// of include/ebiv_gpu.h.
//
// The host keeps what the reference keeps on the host side of Codec.Decode:
// the frame header (ebiv/format.go:376-409), the golden-override split
// (ebiv/codec.go:162-173), the coded header, table and tile framing
// (ebiv/frame.go:197-274, ebiv/rans.go:396-481) and the codec state machine
// (ebiv/codec.go:70-207).  Everything that touches pixels or rANS streams runs
// on the device (ebiv_kernels.hip).
//
// Sessions and steps.  A session is one stream's codec state.  A step takes
// queued packets of any number of sessions (up to a window of frames per
// session), parses their headers and tables on the host in stream order,
// then runs one entropy launch over every tile of every payload, and
// reconstruction in rounds: round r reconstructs the r-th payload of every
// session.  Frame buffers are assigned as if every payload decodes; the tile
// statuses read back at the end of the step say whether that held, and where
// it did not, the session's later payloads are reconstructed again against
// the references Go would have used.  The plain Codec.Decode is a
// one-session, one-frame step.
//
// Frame buffers live in a per-session pool.  A successful decode's buffer
// becomes its ring slot's frame and the next reference, so the per-frame
// visible copy (frame.go:67-77), the work/ref swap (codec.go:189) and the
// per-GOP golden copy (codec.go:183) become pointer moves.  Go's observable
// state is kept: ring slots are consumed before decoding, a failed decode
// leaves ref and the ring alone, and a failed golden override leaves golden
// partially rewritten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/ebiv_gpu.h"
#include "ebiv_device.h"
#include "ebiv_format.h"
#include "ebiv_host.h"

namespace ebiv {

hipError_t upload_tables(const int32_t dct[3][256], const uint8_t zz[3][256]);
void launch_parse(hipStream_t s, int nitems, int max_tile_w, const FrameJob* jobs, const WorkItem* items);
void launch_recon_inter(hipStream_t s, int ngroups, const FrameJob* jobs, const WorkItem* items, int njobs);
void launch_recon_intra(hipStream_t s, int nitems, const FrameJob* jobs, const WorkItem* items, int waves, bool key,
                        bool cells, int band_rows);
void launch_rgba(hipStream_t s, const RgbaJob& j);
void launch_rgba_frames(hipStream_t s, const RgbaJob* jobs, int njobs, uint32_t max_px8);

namespace {

// ---- static tables (transform.go:27-40, scan.go:17-35, quant.go:19-29) ------
struct StaticTables {
  int32_t dct[3][256];
  uint8_t zz[3][256];
  StaticTables() {
    std::memset(dct, 0, sizeof dct);
    std::memset(zz, 0, sizeof zz);
    for (int s = 0; s < 3; s++) {
      const int n = 4 << s;
      int k = 0;
      for (int d = 0; d < 2 * n - 1; d++) {
        if ((d & 1) == 0) {
          for (int y = std::min(d, n - 1); y >= 0 && d - y < n; y--) zz[s][k++] = static_cast<uint8_t>(y * n + d - y);
        } else {
          for (int x = std::min(d, n - 1); x >= 0 && d - x < n; x--) zz[s][k++] = static_cast<uint8_t>((d - x) * n + x);
        }
      }
      for (int kk = 0; kk < n; kk++) {
        const double alpha = kk == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
        for (int i = 0; i < n; i++)
          dct[s][kk * n + i] = static_cast<int32_t>(std::lround(alpha * std::cos((2.0 * i + 1.0) * kk * M_PI / (2.0 * n)) * 8192.0));
      }
    }
  }
};
const StaticTables& statics() {
  static StaticTables t;
  return t;
}

std::mutex g_init_mu;
std::vector<bool> g_dev_ready;

// Staging of ebiv_frame_rgba_noalloc: RGBA of one frame of at most
// kRgbaScratchBytes, converted on the device and copied to the caller's host
// memory.  A file-scope device array, so that entry point allocates nothing:
// each device holds its own copy (in the library's code object) until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory: its device address is taken once per device in
// ensure_device (g_scratch) and used everywhere instead.
constexpr size_t kRgbaScratchBytes = static_cast<size_t>(3840) * 2160 * 4;
static __device__ __align__(256) uint8_t scratch[kRgbaScratchBytes];
std::vector<uint8_t*> g_scratch;  // per device: scratch's device address

int ensure_device(int dev) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) return EBIV_ERR_DEVICE;
  if (static_cast<int>(g_dev_ready.size()) < count) {
    g_dev_ready.resize(count, false);
    g_scratch.resize(count, nullptr);
  }
  if (hipSetDevice(dev) != hipSuccess) return EBIV_ERR_DEVICE;
  if (!g_dev_ready[dev]) {
    const StaticTables& t = statics();
    if (upload_tables(t.dct, t.zz) != hipSuccess) return EBIV_ERR_DEVICE;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EBIV_ERR_DEVICE;
    g_scratch[dev] = static_cast<uint8_t*>(p);
    g_dev_ready[dev] = true;
  }
  return EBIV_OK;
}

inline size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

// Blocks of device memory (per device) and pinned host memory that sessions
// released, kept by size for the next session of the same geometry.  A
// serving loop opens and closes sessions all the time; each new session
// would otherwise pay hipMalloc / hipHostMalloc for its frame pool and
// payload buffers inside the step that decodes its first frame, and each
// closed one hipFree, which waits for the device.  Released blocks are kept
// up to a cap (beyond it they are freed) and held until the process ends.
class BlockCache {
 public:
  void* get(bool pinned, size_t bytes) {
    const int dev = pinned ? -1 : current_device();
    {
      std::lock_guard<std::mutex> g(m_);
      auto it = blocks_.find(Key{dev, bytes});
      if (it != blocks_.end()) {
        void* p = it->second;
        blocks_.erase(it);
        (pinned ? pinned_bytes_ : device_bytes_) -= bytes;
        return p;
      }
    }
    void* p = nullptr;
    const hipError_t rc = pinned ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
    return rc == hipSuccess ? p : nullptr;
  }
  void put(bool pinned, void* p, size_t bytes) {
    if (!p) return;
    const int dev = pinned ? -1 : current_device();
    {
      std::lock_guard<std::mutex> g(m_);
      size_t& held = pinned ? pinned_bytes_ : device_bytes_;
      if (held + bytes <= (pinned ? kMaxPinned : kMaxDevice)) {
        blocks_.emplace(Key{dev, bytes}, p);
        held += bytes;
        return;
      }
    }
    (void)(pinned ? hipHostFree(p) : hipFree(p));
  }

 private:
  static constexpr size_t kMaxDevice = size_t(8) << 30;  // per process, all devices
  static constexpr size_t kMaxPinned = size_t(1) << 30;
  struct Key {
    int dev;
    size_t bytes;
    bool operator<(const Key& o) const { return dev != o.dev ? dev < o.dev : bytes < o.bytes; }
  };
  static int current_device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
  }
  std::mutex m_;
  std::multimap<Key, void*> blocks_;
  size_t device_bytes_ = 0, pinned_bytes_ = 0;
};

BlockCache& block_cache() {
  static BlockCache* c = new BlockCache();  // never destroyed: its blocks live until the process ends
  return *c;
}

template <typename T>
struct DevArray {
  T* d = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    release();
    d = static_cast<T*>(block_cache().get(false, std::max<size_t>(n, 1) * sizeof(T)));
    if (!d) return false;
    cap = n;
    return true;
  }
  void release() {
    if (d) block_cache().put(false, d, std::max<size_t>(cap, 1) * sizeof(T));
    d = nullptr;
    cap = 0;
  }
  ~DevArray() { release(); }
};

template <typename T>
struct PinnedArray {
  T* h = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    release();
    h = static_cast<T*>(block_cache().get(true, std::max<size_t>(n, 1) * sizeof(T)));
    if (!h) return false;
    cap = n;
    return true;
  }
  void release() {
    if (h) block_cache().put(true, h, std::max<size_t>(cap, 1) * sizeof(T));
    h = nullptr;
    cap = 0;
  }
  ~PinnedArray() { release(); }
};

struct DevBuf {
  uint8_t* base = nullptr;
  size_t bytes = 0;
  Planes p{};
};

// Device buffers of one coded payload in flight: the entropy kernel's
// macroblock records and coefficients, the intra residuals, the tile
// statuses; and where the step staged its descriptor (HostPayload::desc).
struct PayloadSlot {
  const uint8_t* hdesc = nullptr;  // the host descriptor (HostCodec::payloads)
  size_t desc_bytes = 0;
  size_t desc_off = 0;  // offset in the step's staging buffer
  int ntiles = 0;
  int max_tile_w = 0;
  bool inter = false;
  bool generic = false;  // no sign table: the entropy kernel runs its generic decoder
  DevArray<MBRec> recs;
  DevArray<uint32_t> coefs;
  DevArray<int16_t> res;
  size_t status_off = 0;  // first tile's entry in the step's status array
  PinnedArray<TileStatus> h_status;  // the last step's tile statuses
};
}  // namespace

// One frame of a session within a step: its host state and the device
// buffers it reads and writes.
struct StepFrame : HostFrame {
  const uint8_t* dev = nullptr;  // packet bytes on the device
  size_t staged_off = 0;
  bool staged = false;
  int64_t ts = 0;
  uint64_t tag = 0;
  // buffer assignment
  int target = -1, gtarget = -1, ref_in = -1, gold_in = -1;
  int out_buf = -1;
};

// One stream's codec: the host state machine (HostCodec) plus its device
// frame pool, payload buffers and host ring.
struct Session : HostCodec {
  int id = 0;
  hipStream_t stream = nullptr;  // the engine's stream (frame pool clears)
  std::vector<DevBuf> pool;
  std::vector<int> ring;  // pool index per slot, -1 = never written
  int ref = -1, golden = -1;
  std::vector<uint8_t*> host_ring;
  std::vector<std::unique_ptr<PayloadSlot>> slots;
  ebiv_stats stats{0, 0, 0, 0};
  // presentation surfaces (ebiv_batch_set_surfaces)
  std::vector<uint8_t*> surfaces;
  size_t surface_pitch = 0;
  int next_surface = 0;
  struct Queued {
    std::vector<uint8_t> host;
    const uint8_t* dev = nullptr;
    int64_t ts = 0;
    uint64_t tag = 0;
  };
  std::deque<Queued> queue;

  ~Session() override {
    release_pool();
    for (auto* h : host_ring) if (h) (void)hipHostFree(h);
  }

  void release_pool() {
    for (auto& b : pool) if (b.base) block_cache().put(false, b.base, b.bytes);
    pool.clear();
  }

  // The frame pool of configure (codec.go:129-153).  Live at a step's
  // start: ring images, ref and golden.  Written by a step: one buffer per
  // frame of the window and a golden-override target.  A failed allocation
  // frees what it took, and the next key frame tries again.
  int allocate_frames() override {
    release_pool();
    const int pw = mbc * kMB, ph = mbr * kMB, cw = mbc * kCMB, ch = mbr * kCMB;
    const size_t ysz = static_cast<size_t>(pw) * ph, csz = align256(static_cast<size_t>(cw) * ch);
    const size_t npool = static_cast<size_t>(ring_size) + 2 + static_cast<size_t>(window) + 1;
    pool.assign(npool, DevBuf{});
    for (auto& b : pool) {
      b.bytes = ysz + 2 * csz;
      b.base = static_cast<uint8_t*>(block_cache().get(false, b.bytes));
      if (!b.base) {
        release_pool();
        return fail(EBIV_ERR_DEVICE, "frame pool allocation failed");
      }
      b.p = Planes{b.base, b.base + ysz, b.base + ysz + csz, pw, ph, cw, ch};
      // (the clear is ordered before every use of the pool: reconstruction,
      // copies and presentation all run on this stream or join it)
      if (hipMemsetAsync(b.base, 0, b.bytes, stream) != hipSuccess) {
        release_pool();
        return fail(EBIV_ERR_DEVICE, "frame pool clear failed");
      }
    }
    ring.assign(static_cast<size_t>(ring_size), -1);
    ref = 0;
    golden = 1;
    return EBIV_OK;
  }

  int ensure_host_ring() {
    if (!host_ring.empty()) return EBIV_OK;
    host_ring.assign(static_cast<size_t>(ring_size), nullptr);
    const size_t hsz = static_cast<size_t>(ystride) * H + 2 * static_cast<size_t>(cstride) * CH;
    for (auto*& hr : host_ring) {
      if (hipHostMalloc(reinterpret_cast<void**>(&hr), hsz, hipHostMallocDefault) != hipSuccess) {
        hr = nullptr;
        return fail(EBIV_ERR_DEVICE, "host ring allocation failed");
      }
      std::memset(hr, 0, hsz);
    }
    return EBIV_OK;
  }

  PayloadSlot* slot(int i) {
    while (static_cast<int>(slots.size()) <= i) slots.emplace_back(new PayloadSlot());
    return slots[static_cast<size_t>(i)].get();
  }

  // Device buffers of payload slot si, sized for this geometry.
  int reserve_payload(int si) override {
    PayloadSlot* ps = slot(si);
    const HostPayload& hp = payloads[static_cast<size_t>(si)];
    ps->hdesc = hp.desc.data();
    ps->desc_bytes = hp.desc.size();
    ps->ntiles = hp.ntiles;
    ps->max_tile_w = hp.max_tile_w;
    ps->inter = hp.inter;
    // (the entropy kernel's own choice, parse_item)
    ps->generic = reinterpret_cast<const FrameHdr*>(hp.desc.data())->vt[kSignVec][kSignLane] == kNoSym;
    const size_t nmb = static_cast<size_t>(mbc) * mbr;
    if (!ps->recs.reserve(nmb) || !ps->coefs.reserve(nmb * kCoefPerMB) ||
        !ps->res.reserve(nmb * (kResPerMB + kResSyncPerMB)) ||
        !ps->h_status.reserve(static_cast<size_t>(hp.ntiles)))
      return fail(EBIV_ERR_DEVICE, "payload buffer allocation failed");
    return EBIV_OK;
  }
  // First failing tile in tile order (decode_intra.go:86-93).
  int tile_result(const PayloadSlot* ps) {
    for (int i = 0; i < ps->ntiles; i++) {
      stats.symbols += ps->h_status.h[i].nsym;
      stats.tiles++;
      if (ps->generic) stats.tiles_generic++;
    }
    for (int i = 0; i < ps->ntiles; i++) {
      const TileStatus& s = ps->h_status.h[i];
      if (s.err) {
        switch (s.err >> 8) {
          case kTileShort: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: tile stream too short to initialize rANS");
          case kTileUnused: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: symbol decoded from unused context");
          default: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: rANS stream underrun");
        }
      }
    }
    return EBIV_OK;
  }

};


// Launch shapes of the reconstruction kernels (see ebiv_batch_set_schedule).
struct Schedule {
  int intra_waves_key = 16;   // wavefronts per key-frame tile, launches of few tiles
  int intra_waves_key_many = 4;  // ... launches of at least many_tiles tiles
  int intra_waves_few = 16;   // wavefronts per inter-frame tile, launches of few tiles
  int intra_waves_many = 4;   // ... launches of at least many_tiles tiles
  int many_tiles = 256;
  int recon_lanes = 1;        // HIP streams the sessions' reconstruction is split over
  bool pipeline = true;       // next step's entropy launch beside this step's reconstruction
  bool intra_cells = false;   // key-frame and intra-dense tiles by block levels (measured slower)
  int key_band_rows = 4;      // macroblock rows per workgroup of a key-frame tile (key_bands)
  // Band rows of a step's key-frame tiles (key_tiles of them), or of its
  // inter-frame tiles (which band only when they turn out all intra): none
  // on the block-level path (whole tiles), nor for many_tiles tiles or more,
  // which fill the chip by themselves (64 x 12 1080p key tiles: 132.7 ms of
  // intra kernel time whole, 155.2 ms in bands of 4).
  int band_rows(size_t tiles) const {
    return !intra_cells && tiles < static_cast<size_t>(many_tiles) ? key_band_rows : 0;
  }
  int inter_intra_waves(size_t ntiles) const {
    return ntiles >= static_cast<size_t>(many_tiles) ? intra_waves_many : intra_waves_few;
  }
  // (a launch of many key-frame tiles keeps more of them resident at once;
  // a few tiles, as in the plain codec, finish sooner with more wavefronts)
  int key_intra_waves(size_t ntiles) const {
    return ntiles >= static_cast<size_t>(many_tiles) ? intra_waves_key_many : intra_waves_key;
  }
};

// Host worker threads for per-session work of a step (header, table and
// tile-framing parse of every queued packet).  parallel_for(n, fn) runs
// fn(0..n-1) on the workers and the calling thread and returns when all
// are done; each worker makes the engine's device current first.
class HostPool {
 public:
  explicit HostPool(int device) : device_(device) {
    const unsigned hc = std::thread::hardware_concurrency();
    const int n = static_cast<int>(hc > 1 ? (hc < 8 ? hc : 8) : 1) - 1;
    for (int i = 0; i < n; i++) workers_.emplace_back([this] { run(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      quit_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  void parallel_for(int n, const std::function<void(int)>& fn) {
    if (n <= 1 || workers_.empty()) {
      for (int i = 0; i < n; i++) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      active_ = static_cast<int>(workers_.size());
      gen_++;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> l(mu_);
    done_cv_.wait(l, [this] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void drain() {
    for (int i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*fn_)(i);
  }
  void run() {
    (void)hipSetDevice(device_);
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return quit_ || gen_ != seen; });
        if (quit_) return;
        seen = gen_;
      }
      drain();
      {
        std::lock_guard<std::mutex> g(mu_);
        if (--active_ == 0) done_cv_.notify_one();
      }
    }
  }
  int device_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  int n_ = 0, active_ = 0;
  std::atomic<int> next_{0};
  uint64_t gen_ = 0;
  bool quit_ = false;
};

struct JobRef {
  size_t si, fi;
  bool ov;
};

// One step of the engine in flight: the frames it takes from each session,
// their packets, the staging and job tables of its launches and its tile
// statuses.  The engine keeps two, so that one step's entropy launch can run
// while the previous step reconstructs.
struct StepBuf {
  std::vector<Session*> ss;  // null: a session closed while the step was in flight
  std::vector<std::vector<StepFrame>> frames;
  std::vector<std::vector<Session::Queued>> held;  // the frames' packets
  std::vector<JobRef> jrefs;
  size_t nparse = 0;
  int max_tw = 1;
  // One staging blob, host and device, uploaded in as few copies as the step
  // allows: host packets and descriptors, then the job table (hj/dj), then
  // the work items (hi/di: entropy items, then reconstruction items).
  PinnedArray<uint8_t> h_stage;
  DevArray<uint8_t> d_stage;
  size_t job_off = 0, item_off = 0;
  int band_rows = 0;        // key-frame band rows of this step's launches (Schedule::band_rows)
  int band_rows_inter = 0;  // the same for its inter-frame tiles
  bool cells = false;   // this step's intra launches take the block-level path
  FrameJob* hj = nullptr;
  FrameJob* dj = nullptr;
  WorkItem* hi = nullptr;
  WorkItem* di = nullptr;
  // The reconstruction plan (plan_recon): rounds of jobs, and per round and
  // lane q the item ranges of its inter launch, key-frame intra launch and
  // inter-frame intra launch.
  bool planned = false;
  int lanes = 1;
  size_t items_end = 0;
  std::vector<std::vector<size_t>> rounds;
  std::vector<size_t> inter_off, intra_off, key_end, dense_off, round_end;
  std::vector<uint32_t> inter_groups_of;
  std::vector<uint64_t> inter_mbs_of;
  DevArray<TileStatus> d_status;  // every payload's tiles, contiguous: one download
  PinnedArray<TileStatus> h_status;
  hipEvent_t parsed = nullptr;  // the statuses are on the host
  std::vector<std::vector<char>> okm, oko;  // per frame: its payloads' tiles all decoded
  bool live = false;  // entropy launched, step not yet returned

  ~StepBuf() {
    if (parsed) (void)hipEventDestroy(parsed);
  }
  PayloadSlot* payload(const JobRef& r) {
    const StepFrame& f = frames[r.si][r.fi];
    return ss[r.si]->slot(r.ov ? f.ov_slot : f.main_slot);
  }
  void clear() {
    ss.clear();
    frames.clear();
    held.clear();
    jrefs.clear();
    okm.clear();
    oko.clear();
    nparse = 0;
    max_tw = 1;
    live = false;
    planned = false;
    items_end = 0;
  }
  size_t frames_of(int session) const {
    size_t n = 0;
    for (size_t i = 0; i < ss.size(); i++)
      if (ss[i] && (session == 0 || ss[i]->id == session)) n += frames[i].size();
    return n;
  }
};

struct Engine {
  int device = 0;
  Schedule sched;
  hipStream_t stream = nullptr;   // reconstruction (lane 0), and everything in the plain codec
  hipStream_t pstream = nullptr;  // the batch engine's entropy launches
  std::vector<std::unique_ptr<Session>> sessions;
  int next_id = 1;
  size_t rotate = 0;  // first session served by the next batch step
  StepBuf steps[2];
  int cur = 0;  // the step whose entropy launch is in flight (or is taken next)
  std::vector<ebiv_completion> carry;  // completions beyond a call's max_out
  DevArray<uint32_t> d_rgba;
  PinnedArray<RgbaJob> h_present;  // a step's presentation conversions
  DevArray<RgbaJob> d_present;
  std::string msg;
  bool timing = false;
  // Reconstruction lanes beyond `stream` (default priority, created on first
  // use) and the events that join them.
  static constexpr int kMaxLanes = 4;
  hipStream_t lanes[kMaxLanes - 1] = {};
  hipEvent_t lane_fork = nullptr, lane_join[kMaxLanes - 1] = {};
  struct Mark {
    int kind;
    uint64_t mbs;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;      // timed launches not yet collected
  std::vector<hipEvent_t> spare;  // events of collected marks
  ebiv_kernel_times ktimes{};

  std::unique_ptr<HostPool> pool;  // created on the first batch step

  ~Engine() {
    pool.reset();
    if (pstream) (void)hipStreamSynchronize(pstream);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto l : lanes) if (l) (void)hipStreamSynchronize(l);
    sessions.clear();
    for (auto& m : marks) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    for (auto e : spare) (void)hipEventDestroy(e);
    for (int i = 0; i < kMaxLanes - 1; i++) {
      if (lane_join[i]) (void)hipEventDestroy(lane_join[i]);
      if (lanes[i]) (void)hipStreamDestroy(lanes[i]);
    }
    if (lane_fork) (void)hipEventDestroy(lane_fork);
    if (pstream) (void)hipStreamDestroy(pstream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int fail(int code, const std::string& m) {
    msg = m;
    return code;
  }

  Session* find(int id) {
    for (auto& s : sessions) if (s->id == id) return s.get();
    return nullptr;
  }

  // Streams and events of `n` reconstruction lanes (lane 0 is `stream`).
  bool ensure_lanes(int n) {
    if (!lane_fork && hipEventCreateWithFlags(&lane_fork, hipEventDisableTiming) != hipSuccess) return false;
    for (int i = 0; i + 1 < n; i++) {
      if (!lanes[i] && hipStreamCreateWithFlags(&lanes[i], hipStreamNonBlocking) != hipSuccess) return false;
      if (!lane_join[i] && hipEventCreateWithFlags(&lane_join[i], hipEventDisableTiming) != hipSuccess) return false;
    }
    return true;
  }
  hipStream_t lane_stream(int g) const { return g == 0 ? stream : lanes[g - 1]; }

  // Brackets a launch of kind k (0 parse, 1 inter, 2 intra, 3 rgba) on
  // stream st with events while timing is on; returns the mark or -1.
  // `mbs`: the macroblocks an inter launch covers.
  hipEvent_t take_event() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
  }
  int mark_begin(int kind, uint64_t mbs = 0, hipStream_t st = nullptr) {
    if (!timing) return -1;
    Mark m{kind, mbs, take_event(), take_event()};
    if (!m.a || !m.b || hipEventRecord(m.a, st ? st : stream) != hipSuccess) {
      if (m.a) spare.push_back(m.a);
      if (m.b) spare.push_back(m.b);
      return -1;
    }
    marks.push_back(m);
    return static_cast<int>(marks.size()) - 1;
  }
  void mark_end(int m, hipStream_t st = nullptr) {
    if (m >= 0) (void)hipEventRecord(marks[static_cast<size_t>(m)].b, st ? st : stream);
  }
  // Adds the timed launches that have finished; the others stay pending.
  void collect_marks() {
    std::vector<Mark> keep;
    for (const Mark& m : marks) {
      if (hipEventQuery(m.b) != hipSuccess) {
        keep.push_back(m);
        continue;
      }
      float ms = 0;
      if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
        switch (m.kind) {
          case 0: ktimes.parse_ms += ms; ktimes.parse_n++; break;
          case 1: ktimes.inter_ms += ms; ktimes.inter_n++; ktimes.inter_mbs += m.mbs; break;
          case 2: ktimes.intra_ms += ms; ktimes.intra_n++; break;
          default: ktimes.rgba_ms += ms; ktimes.rgba_n++; break;
        }
      }
      spare.push_back(m.a);
      spare.push_back(m.b);
    }
    marks.swap(keep);
  }

  // Phases of a step (see run_plain and the batch step).
  int launch_entropy(StepBuf& S, hipStream_t ps, bool speculative);
  int download_status(StepBuf& S, hipStream_t st);
  int finish_entropy(StepBuf& S);
  int plan_recon(StepBuf& S);
  int launch_recon(StepBuf& S);
  void commit(StepBuf& S);
  int wait_recon();
  int run_plain(StepBuf& S);
};

namespace {

// Assigns a session's frames their target buffers and inputs by walking them
// in stream order as Go's state machine does (codec.go:174-190).  ok_main /
// ok_ov are each payload's outcome.  Targets are never buffers live at the
// step's start, nor each other.
bool assign_buffers(const Session& s, std::vector<StepFrame>& fr, const std::vector<char>& ok_main,
                    const std::vector<char>& ok_ov) {
  std::vector<int> taken;
  auto live = [&](int b) {
    if (b == s.ref || b == s.golden) return true;
    for (int r : s.ring) if (r == b) return true;
    for (int t : taken) if (t == b) return true;
    return false;
  };
  auto pick = [&]() {
    for (int i = 0; i < static_cast<int>(s.pool.size()); i++)
      if (!live(i)) { taken.push_back(i); return i; }
    return -1;
  };
  int ref = s.ref, golden = s.golden;
  for (size_t k = 0; k < fr.size(); k++) {
    StepFrame& f = fr[k];
    if (f.rc || !f.device_work) continue;
    if ((f.target = pick()) < 0) return false;
    if (f.raw) continue;  // a raw frame writes its ring image only
    f.ref_in = ref;
    f.gold_in = golden;
    if (f.ov_slot >= 0 && (f.gtarget = pick()) < 0) return false;
    if (!ok_main[k]) continue;
    if (f.key) {
      if (f.has_ov) {
        if (f.ov_rc) continue;
        golden = f.gtarget;  // decoded, or partially rewritten on failure
        if (!ok_ov[k]) continue;
      } else {
        golden = f.target;
      }
    }
    ref = f.target;
  }
  return true;
}

// Whether a payload is reconstructed: its frame's own payload when its tiles
// decoded; an override when its key decoded (a failing override still
// rewrites golden up to where it failed, codec.go:176-181).
bool reconstructs(const StepBuf& S, const JobRef& r) { return S.okm[r.si][r.fi] != 0; }

}  // namespace

// Stages the step's packets and descriptors, builds its job table and
// entropy items and launches the entropy kernel over every tile of every
// payload on `ps`.  `speculative` (a one-frame step): the reconstruction's
// frame buffers are assigned as if every payload decodes, which is exact for
// a lone frame, and the reconstruction is planned here too, so that one copy
// uploads everything and the caller downloads the statuses after the
// reconstruction; otherwise the statuses follow the entropy launch and the
// reconstruction is planned from them.
int Engine::launch_entropy(StepBuf& S, hipStream_t ps, bool speculative) {
  auto& ss = S.ss;
  auto& frames = S.frames;
  // ---- the blob's layout: staged packets and descriptors, jobs, items
  size_t stage = 0;
  for (size_t i = 0; i < ss.size(); i++)
    for (auto& f : frames[i]) {
      if (f.rc || !f.device_work) continue;
      if (!f.dev) {
        f.staged_off = stage;
        f.staged = true;
        stage = align256(stage + f.len + 256);  // the entropy kernel reads 256-byte windows
      }
      for (int sl : {f.main_slot, f.ov_slot}) {
        if (sl < 0) continue;
        PayloadSlot* p = ss[i]->slot(sl);
        p->desc_off = stage;
        stage = align256(stage + p->desc_bytes);
      }
    }
  S.jrefs.clear();
  for (size_t i = 0; i < ss.size(); i++)
    for (size_t k = 0; k < frames[i].size(); k++) {
      const StepFrame& f = frames[i][k];
      if (f.rc || !f.device_work || f.raw) continue;
      S.jrefs.push_back({i, k, false});
      if (f.ov_slot >= 0) S.jrefs.push_back({i, k, true});
    }
  const size_t njobs = S.jrefs.size();
  size_t nitems = 0, nparse = 0;
  int max_tw = 1;
  uint32_t max_tile_bytes = 1;
  auto tiles_of = [](const PayloadSlot* p) { return reinterpret_cast<const TileDesc*>(p->hdesc + desc_tiles_off()); };
  // (fixed for the step: the band rows set the item count)
  {
    size_t key_tiles = 0, inter_tiles = 0;
    for (const JobRef& jr : S.jrefs)
      (S.payload(jr)->inter ? inter_tiles : key_tiles) += static_cast<size_t>(S.payload(jr)->ntiles);
    S.band_rows = sched.band_rows(key_tiles);
    S.band_rows_inter = sched.band_rows(inter_tiles);
  }
  S.cells = sched.intra_cells;
  for (const JobRef& jr : S.jrefs) {
    PayloadSlot* p = S.payload(jr);
    p->status_off = nparse;
    nparse += static_cast<size_t>(p->ntiles);
    nitems += static_cast<size_t>(p->ntiles) + 1;  // entropy items, the inter item
    max_tw = std::max(max_tw, p->max_tile_w);
    for (int t = 0; t < p->ntiles; t++) {
      const TileDesc& td = tiles_of(p)[t];
      max_tile_bytes = std::max(max_tile_bytes, td.byte_len);
      nitems += static_cast<size_t>(key_bands(td.x1 - td.x0, td.y1 - td.y0,
                                              p->inter ? S.band_rows_inter : S.band_rows));
    }
  }
  S.nparse = nparse;
  S.max_tw = max_tw;
  S.job_off = align256(stage);
  S.item_off = align256(S.job_off + sizeof(FrameJob) * (njobs + 1));
  const size_t blob = S.item_off + sizeof(WorkItem) * (nitems + 1);
  // (reserve keeps no contents: before anything is written)
  if (!S.h_stage.reserve(blob) || !S.d_stage.reserve(blob) || !S.d_status.reserve(nparse + 1) ||
      !S.h_status.reserve(nparse + 1))
    return fail(EBIV_ERR_DEVICE, "staging allocation failed");
  S.hj = reinterpret_cast<FrameJob*>(S.h_stage.h + S.job_off);
  S.dj = reinterpret_cast<FrameJob*>(S.d_stage.d + S.job_off);
  S.hi = reinterpret_cast<WorkItem*>(S.h_stage.h + S.item_off);
  S.di = reinterpret_cast<WorkItem*>(S.d_stage.d + S.item_off);
  for (size_t i = 0; i < ss.size(); i++)
    for (auto& f : frames[i]) {
      if (f.rc || !f.device_work) continue;
      if (f.staged) {
        std::memcpy(S.h_stage.h + f.staged_off, f.host, f.len);
        f.dev = S.d_stage.d + f.staged_off;
      }
      for (int sl : {f.main_slot, f.ov_slot}) {
        if (sl < 0) continue;
        PayloadSlot* p = ss[i]->slot(sl);
        std::memcpy(S.h_stage.h + p->desc_off, p->hdesc, p->desc_bytes);
      }
    }
  if (speculative)
    for (size_t i = 0; i < ss.size(); i++) {
      const std::vector<char> ok(frames[i].size(), 1);
      if (!assign_buffers(*ss[i], frames[i], ok, ok)) return fail(EBIV_ERR_DEVICE, "frame pool exhausted");
    }

  // ---- jobs: one per coded payload, a session's in stream order
  for (size_t j = 0; j < njobs; j++) {
    const JobRef& r = S.jrefs[j];
    const Session& s = *ss[r.si];
    const StepFrame& f = frames[r.si][r.fi];
    PayloadSlot* p = S.payload(r);
    FrameJob& jb = S.hj[j];
    std::memset(&jb, 0, sizeof jb);
    jb.pkt = f.dev;
    jb.desc = S.d_stage.d + p->desc_off;
    jb.recs = p->recs.d;
    jb.coefs = p->coefs.d;
    jb.res = p->res.d;
    jb.status = S.d_status.d + p->status_off;
    jb.mbc = s.mbc;
    jb.mbr = s.mbr;
  }
  // Entropy items: one per tile, with an issue priority in bits 30-31 of
  // the tile index.  The launch lasts as long as its longest serial tile
  // chain, and the chains of the biggest tiles (key frames) would otherwise
  // share their SIMD's issue slots evenly with many short ones; the biggest
  // get priority 3, those above 1/4 and 1/8 of the biggest 2 and 1.  Tiles go
  // in order of decreasing coded size: wavefronts are dispatched in block
  // order, so the longest chains start first and the short ones fill in
  // behind them.
  std::vector<std::pair<uint32_t, WorkItem>> order;
  order.reserve(nparse);
  for (size_t j = 0; j < njobs; j++) {
    const PayloadSlot* p = S.payload(S.jrefs[j]);
    for (int t = 0; t < p->ntiles; t++) {
      const uint32_t len = tiles_of(p)[t].byte_len;
      const uint64_t l = len;
      const uint32_t prio = 2 * l >= max_tile_bytes ? 3 : 4 * l >= max_tile_bytes ? 2 : 8 * l >= max_tile_bytes ? 1 : 0;
      order.emplace_back(len, WorkItem{static_cast<uint32_t>(j), static_cast<uint32_t>(t) | prio << 30});
    }
  }
  std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  for (size_t w = 0; w < order.size(); w++) S.hi[w] = order[w].second;
  if (speculative) {  // the frame buffers assigned above, then the reconstruction's items
    for (size_t j = 0; j < njobs; j++) {
      const JobRef& r = S.jrefs[j];
      const Session& s = *ss[r.si];
      const StepFrame& f = frames[r.si][r.fi];
      FrameJob& jb = S.hj[j];
      jb.rec = s.pool[static_cast<size_t>(r.ov ? f.gtarget : f.target)].p;
      jb.ref = s.pool[static_cast<size_t>(r.ov ? f.target : f.ref_in)].p;
      jb.gold = s.pool[static_cast<size_t>(r.ov ? f.target : f.gold_in)].p;
    }
    if (int rc = plan_recon(S)) return rc;
  }
  const size_t used = S.item_off + sizeof(WorkItem) * (speculative ? S.items_end : nparse);
  if (hipMemcpyAsync(S.d_stage.d, S.h_stage.h, used, hipMemcpyHostToDevice, ps) != hipSuccess)
    return fail(EBIV_ERR_DEVICE, "staging upload failed");
  if (nparse) {
    const int m = mark_begin(0, 0, ps);
    launch_parse(ps, static_cast<int>(nparse), max_tw, S.dj, S.di);
    mark_end(m, ps);
  }
  if (hipGetLastError() != hipSuccess) return fail(EBIV_ERR_DEVICE, "kernel launch failed");
  if (!speculative)
    if (int rc = download_status(S, ps)) return rc;
  S.live = true;
  return EBIV_OK;
}

// The step's tile statuses to the host on `st`, then the event
// finish_entropy waits for.
int Engine::download_status(StepBuf& S, hipStream_t st) {
  if (S.nparse && hipMemcpyAsync(S.h_status.h, S.d_status.d, sizeof(TileStatus) * S.nparse, hipMemcpyDeviceToHost,
                                 st) != hipSuccess)
    return fail(EBIV_ERR_DEVICE, "status download failed");
  if (!S.parsed && hipEventCreateWithFlags(&S.parsed, hipEventDisableTiming) != hipSuccess) {
    S.parsed = nullptr;
    return fail(EBIV_ERR_DEVICE, "event creation failed");
  }
  if (hipEventRecord(S.parsed, st) != hipSuccess) return fail(EBIV_ERR_DEVICE, "event record failed");
  return EBIV_OK;
}

// Waits for the step's tile statuses and turns them into each payload's
// outcome (the first failing tile in tile order, decode_intra.go:86-93).
int Engine::finish_entropy(StepBuf& S) {
  if (hipEventSynchronize(S.parsed) != hipSuccess) return fail(EBIV_ERR_DEVICE, "decode failed on the device");
  S.okm.assign(S.ss.size(), {});
  S.oko.assign(S.ss.size(), {});
  for (size_t i = 0; i < S.ss.size(); i++) {
    S.okm[i].assign(S.frames[i].size(), 1);
    S.oko[i].assign(S.frames[i].size(), 1);
  }
  for (const JobRef& jr : S.jrefs) {
    if (!S.ss[jr.si]) continue;
    Session& s = *S.ss[jr.si];
    StepFrame& f = S.frames[jr.si][jr.fi];
    PayloadSlot* p = S.payload(jr);
    std::memcpy(p->h_status.h, S.h_status.h + p->status_off, sizeof(TileStatus) * static_cast<size_t>(p->ntiles));
    if (s.tile_result(p) == EBIV_OK) continue;
    if (jr.ov) {
      S.oko[jr.si][jr.fi] = 0;
      f.ov_msg = s.msg;
    } else {
      S.okm[jr.si][jr.fi] = 0;
      f.msg = s.msg;
    }
  }
  return EBIV_OK;
}

// Reconstructs the step.  Buffers are assigned from the payloads' real
// outcomes (unless launch_entropy assigned them for a one-frame step), so
// every payload is reconstructed once, against the references Go uses.
// Runs in rounds: round r reconstructs the r-th reconstructed payload of
// every session.  Per round and lane: inter macroblock groups (key frames
// too: the launch forms their intra residuals), then the intra tiles of
// key-frame payloads (their own launch: many wavefronts per tile and the LDS
// frontier, for latency), then those of inter payloads (few wavefronts per
// tile, for throughput).  Sessions are dealt to lanes round robin; a lane's
// launches run in order on its own stream, and the lanes overlap.  Sessions
// share nothing, so the split leaves every result unchanged.
int Engine::plan_recon(StepBuf& S) {
  auto& ss = S.ss;
  auto& frames = S.frames;
  const bool exact = !S.okm.empty();  // statuses known: buffers from the real outcomes
  if (exact) {
    for (size_t i = 0; i < ss.size(); i++)
      if (ss[i] && !assign_buffers(*ss[i], frames[i], S.okm[i], S.oko[i])) return fail(EBIV_ERR_DEVICE, "frame pool exhausted");
    for (size_t j = 0; j < S.jrefs.size(); j++) {
      const JobRef& r = S.jrefs[j];
      if (!ss[r.si]) continue;
      const Session& s = *ss[r.si];
      const StepFrame& f = frames[r.si][r.fi];
      FrameJob& jb = S.hj[j];
      jb.rec = s.pool[static_cast<size_t>(r.ov ? f.gtarget : f.target)].p;
      jb.ref = s.pool[static_cast<size_t>(r.ov ? f.target : f.ref_in)].p;
      jb.gold = s.pool[static_cast<size_t>(r.ov ? f.target : f.gold_in)].p;
    }
  }
  const int nl = std::max(1, std::min({sched.recon_lanes, static_cast<int>(kMaxLanes), static_cast<int>(ss.size())}));
  S.lanes = nl;
  auto lane_of = [&](size_t si) { return static_cast<int>(si % static_cast<size_t>(nl)); };
  // Rounds by dependency depth: a payload's round is one after the latest
  // round of the payloads of this step whose output it reads (its ref and
  // golden; an override, its key and the golden it copies), and 0 when it
  // reads none.  Intra-coded payloads (key frames) read no reference, so
  // every key frame of the step reconstructs in round 0, beside the others
  // there, instead of each holding up the round of its stream position.
  // Targets are fresh buffers, so no payload writes what another still reads.
  auto& rounds = S.rounds;
  rounds.clear();
  {
    std::vector<int> level(S.jrefs.size(), -1);
    std::vector<int> writer;  // pool buffer -> job writing it in this step
    size_t cur_si = SIZE_MAX;
    for (size_t j = 0; j < S.jrefs.size(); j++) {
      const JobRef& r = S.jrefs[j];
      if (!ss[r.si] || (exact && !reconstructs(S, r))) continue;
      if (r.si != cur_si) {  // (a session's jobs are contiguous, in stream order)
        cur_si = r.si;
        writer.assign(ss[r.si]->pool.size(), -1);
      }
      const StepFrame& f = frames[r.si][r.fi];
      int lv = 0;
      auto dep = [&](int buf) {
        if (buf >= 0 && writer[static_cast<size_t>(buf)] >= 0)
          lv = std::max(lv, level[static_cast<size_t>(writer[static_cast<size_t>(buf)])] + 1);
      };
      if (r.ov) {
        dep(f.target);
        dep(f.gold_in);
      } else if (S.payload(r)->inter) {
        dep(f.ref_in);
        dep(f.gold_in);
      }
      level[j] = lv;
      writer[static_cast<size_t>(r.ov ? f.gtarget : f.target)] = static_cast<int>(j);
      if (rounds.size() <= static_cast<size_t>(lv)) rounds.resize(static_cast<size_t>(lv) + 1);
      rounds[static_cast<size_t>(lv)].push_back(j);
    }
  }
  auto inter_groups = [&](size_t j) -> uint32_t {
    const Session& s = *ss[S.jrefs[j].si];
    return (static_cast<uint32_t>(s.mbc * s.mbr) + kInterGroupMbs - 1) / kInterGroupMbs;
  };
  auto is_key = [&](size_t j) { return !S.payload(S.jrefs[j])->inter; };
  size_t w = S.nparse;
  const size_t nrl = rounds.size() * static_cast<size_t>(nl);
  auto& inter_off = S.inter_off;
  auto& intra_off = S.intra_off;
  auto& key_end = S.key_end;
  auto& dense_off = S.dense_off;
  auto& round_end = S.round_end;
  auto& inter_groups_of = S.inter_groups_of;
  auto& inter_mbs_of = S.inter_mbs_of;
  inter_off.assign(nrl, 0);
  intra_off.assign(nrl, 0);
  key_end.assign(nrl, 0);
  dense_off.assign(nrl, 0);
  round_end.assign(nrl, 0);
  // With the tiles' statuses known and inter-frame tiles unbanded (a step of
  // many tiles), an inter-frame tile coded mostly intra (a scene cut or a
  // flash coded as an inter frame) goes to a launch of its own with the LDS
  // frontier, shaped as a key-frame launch; the kernel takes the frontier
  // path by the same test.  In the launch of the other tiles (no frontier)
  // it would run anti-diagonal rounds and hold up the whole round.
  const bool split = exact && S.band_rows_inter == 0;
  auto dense = [&](size_t j, int t) {
    const TileStatus& st = reinterpret_cast<const TileStatus*>(S.payload(S.jrefs[j])->h_status.h)[t];
    return 2 * st.n_intra > st.mbs_done;
  };
  inter_groups_of.assign(nrl, 0);
  inter_mbs_of.assign(nrl, 0);
  for (size_t r = 0; r < rounds.size(); r++)
    for (int g = 0; g < nl; g++) {
      const size_t q = r * static_cast<size_t>(nl) + static_cast<size_t>(g);
      // Inter launches take one item per job: {job, its first macroblock
      // group in the launch}; the kernel finds a workgroup's job among them.
      inter_off[q] = w;
      uint32_t prefix = 0;
      for (size_t j : rounds[r])
        if (lane_of(S.jrefs[j].si) == g) {
          S.hi[w++] = WorkItem{static_cast<uint32_t>(j), prefix};
          prefix += inter_groups(j);
          if (!is_key(j)) inter_mbs_of[q] += static_cast<uint64_t>(ss[S.jrefs[j].si]->mbc) * ss[S.jrefs[j].si]->mbr;
        }
      inter_groups_of[q] = prefix;
      intra_off[q] = w;
      // key-frame tiles, then inter-frame tiles, then (split) the inter-frame
      // tiles coded mostly intra
      for (int pass = 0; pass < 3; pass++) {
        for (size_t j : rounds[r])
          if (lane_of(S.jrefs[j].si) == g && is_key(j) == (pass == 0))
            for (int t = 0; t < S.payload(S.jrefs[j])->ntiles; t++) {
              if (pass > 0 && (split ? dense(j, t) : false) != (pass == 2)) continue;
              // a banded key-frame tile: one item per band, the same item
              // (each workgroup claims its band on the device)
              const uint8_t* hd = S.payload(S.jrefs[j])->hdesc;
              const TileDesc& td = reinterpret_cast<const TileDesc*>(hd + desc_tiles_off())[t];
              const int nb = key_bands(td.x1 - td.x0, td.y1 - td.y0, pass == 0 ? S.band_rows : S.band_rows_inter);
              for (int b = 0; b < nb; b++) S.hi[w++] = WorkItem{static_cast<uint32_t>(j), static_cast<uint32_t>(t)};
            }
        if (pass == 0) key_end[q] = w;
        if (pass == 1) dense_off[q] = w;
      }
      round_end[q] = w;
    }
  S.items_end = w;
  S.planned = true;
  return EBIV_OK;
}

// Reconstructs the step: plans it (unless launch_entropy did), uploads the
// job table and items the plan changed (one copy, from the job table to the
// last item), and launches the rounds.
int Engine::launch_recon(StepBuf& S) {
  auto& ss = S.ss;
  auto& frames = S.frames;
  if (!S.planned) {
    if (int rc = plan_recon(S)) return rc;
    const size_t from = S.job_off, to = S.item_off + sizeof(WorkItem) * S.items_end;
    if (!S.jrefs.empty() &&
        hipMemcpyAsync(S.d_stage.d + from, S.h_stage.h + from, to - from, hipMemcpyHostToDevice, stream) != hipSuccess)
      return fail(EBIV_ERR_DEVICE, "job table upload failed");
  }
  const int nl = S.lanes;
  if (nl > 1 && !ensure_lanes(nl)) return fail(EBIV_ERR_DEVICE, "stream creation failed");
  auto lane_of = [&](size_t si) { return static_cast<int>(si % static_cast<size_t>(nl)); };
  const auto& rounds = S.rounds;
  const auto& inter_off = S.inter_off;
  const auto& intra_off = S.intra_off;
  const auto& key_end = S.key_end;
  const auto& dense_off = S.dense_off;
  const auto& round_end = S.round_end;
  const auto& inter_groups_of = S.inter_groups_of;
  const auto& inter_mbs_of = S.inter_mbs_of;

  // ---- raw frames: packed planes into their buffers (codec.go:195-207)
  for (size_t i = 0; i < ss.size(); i++) {
    if (!ss[i]) continue;
    for (auto& f : frames[i]) {
      if (f.rc || !f.raw) continue;
      const Session& s = *ss[i];
      const DevBuf& b = s.pool[static_cast<size_t>(f.target)];
      const uint8_t* src = f.dev + f.body_off;
      const size_t ysz = static_cast<size_t>(s.W) * s.H, csz = static_cast<size_t>(s.CW) * s.CH;
      if (hipMemcpy2DAsync(b.p.y, b.p.pw, src, s.W, s.W, s.H, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
          hipMemcpy2DAsync(b.p.cb, b.p.cw, src + ysz, s.CW, s.CW, s.CH, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
          hipMemcpy2DAsync(b.p.cr, b.p.cw, src + ysz + csz, s.CW, s.CW, s.CH, hipMemcpyDeviceToDevice, stream) != hipSuccess)
        return fail(EBIV_ERR_DEVICE, "raw frame copy failed");
    }
  }
  if (nl > 1) {  // the lanes start after the uploads above
    if (hipEventRecord(lane_fork, stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "event record failed");
    for (int g = 1; g < nl; g++)
      if (hipStreamWaitEvent(lanes[g - 1], lane_fork, 0) != hipSuccess) return fail(EBIV_ERR_DEVICE, "stream wait failed");
  }
  // A golden override decodes over a copy of the golden it replaces, so a
  // failing override leaves Go's partial rewrite (codec.go:178-181).
  auto golden_copy = [&](const JobRef& jr, hipStream_t st) -> bool {
    const Session& s = *ss[jr.si];
    const StepFrame& f = frames[jr.si][jr.fi];
    const DevBuf& src = s.pool[static_cast<size_t>(f.gold_in)];
    return hipMemcpyAsync(s.pool[static_cast<size_t>(f.gtarget)].base, src.base, src.bytes, hipMemcpyDeviceToDevice,
                          st) == hipSuccess;
  };
  for (size_t r = 0; r < rounds.size(); r++) {
    for (int g = 0; g < nl; g++) {
      const size_t q = r * static_cast<size_t>(nl) + static_cast<size_t>(g);
      hipStream_t st = lane_stream(g);
      for (size_t j : rounds[r])
        if (lane_of(S.jrefs[j].si) == g && S.jrefs[j].ov && !golden_copy(S.jrefs[j], st))
          return fail(EBIV_ERR_DEVICE, "golden copy failed");
      if (intra_off[q] > inter_off[q]) {
        const int m = mark_begin(1, inter_mbs_of[q], st);
        launch_recon_inter(st, static_cast<int>(inter_groups_of[q]), S.dj, S.di + inter_off[q],
                           static_cast<int>(intra_off[q] - inter_off[q]));
        mark_end(m, st);
      }
      if (key_end[q] > intra_off[q]) {
        const int m = mark_begin(2, 0, st);
        const size_t n = key_end[q] - intra_off[q];
        // (a banded tile's workgroup holds band_rows rows: no more waves)
        const int kw = S.band_rows > 0 ? std::min(sched.key_intra_waves(n), S.band_rows) : sched.key_intra_waves(n);
        launch_recon_intra(st, static_cast<int>(n), S.dj, S.di + intra_off[q], kw, true, S.cells, S.band_rows);
        mark_end(m, st);
      }
      if (dense_off[q] > key_end[q]) {
        const size_t n = dense_off[q] - key_end[q];
        const int m = mark_begin(2, 0, st);
        // (with bands, the frontier for the tiles coded all intra)
        launch_recon_intra(st, static_cast<int>(n), S.dj, S.di + key_end[q], sched.inter_intra_waves(n),
                           S.band_rows_inter > 0, S.cells, S.band_rows_inter);
        mark_end(m, st);
      }
      if (round_end[q] > dense_off[q]) {  // inter-frame tiles coded mostly intra, split off
        // (shaped as a key-frame launch: many wavefronts per tile when few
        // tiles, few when many, so that the tiles' LDS keeps them resident)
        const size_t n = round_end[q] - dense_off[q];
        const int m = mark_begin(2, 0, st);
        launch_recon_intra(st, static_cast<int>(n), S.dj, S.di + dense_off[q], sched.key_intra_waves(n), true, S.cells,
                           0);
        mark_end(m, st);
      }
    }
  }
  for (int g = 1; g < nl; g++)
    if (hipEventRecord(lane_join[g - 1], lanes[g - 1]) != hipSuccess ||
        hipStreamWaitEvent(stream, lane_join[g - 1], 0) != hipSuccess)
      return fail(EBIV_ERR_DEVICE, "lane join failed");
  if (hipGetLastError() != hipSuccess) return fail(EBIV_ERR_DEVICE, "kernel launch failed");
  return EBIV_OK;
}

// Go's state transitions in stream order (codec.go:84-110, 174-190), from
// the payloads' outcomes.
void Engine::commit(StepBuf& S) {
  for (size_t i = 0; i < S.ss.size(); i++) {
    if (!S.ss[i]) continue;
    Session& s = *S.ss[i];
    for (size_t k = 0; k < S.frames[i].size(); k++) {
      StepFrame& f = S.frames[i][k];
      if (f.rc) continue;  // failed on the host: nothing ran on the device
      const int st = s.resolve(f, S.okm[i][k] != 0, S.oko[i][k] != 0);
      if (f.golden_written) s.golden = f.has_ov ? f.gtarget : f.target;
      if (st) {
        f.rc = st;
        continue;
      }
      s.ring[static_cast<size_t>(f.ring_slot)] = f.target;
      if (!f.raw) s.ref = f.target;
      f.out_buf = f.target;
      s.stats.frames++;
    }
  }
}

int Engine::wait_recon() {
  if (hipStreamSynchronize(stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "decode failed on the device");
  collect_marks();
  return EBIV_OK;
}

// A one-frame step of the plain codec: entropy and reconstruction back to
// back on one stream (buffers assigned as if the frame decodes: exact for a
// lone frame, whose failure discards its output, and whose failing golden
// override keeps the partial rewrite the reconstruction made), then one
// synchronisation.
int Engine::run_plain(StepBuf& S) {
  S.okm.clear();
  if (int rc = launch_entropy(S, stream, true)) return rc;
  if (int rc = launch_recon(S)) return rc;
  if (int rc = download_status(S, stream)) return rc;
  if (int rc = wait_recon()) return rc;
  if (int rc = finish_entropy(S)) return rc;
  commit(S);
  S.live = false;
  return EBIV_OK;
}

}  // namespace ebiv

using namespace ebiv;

struct ebiv_codec {
  Engine eng;
  Session* s = nullptr;
};

struct ebiv_batch {
  Engine eng;
};

namespace {

void fill_frame(const Session& s, int buf, int slot, int64_t ts, bool on_device, ebiv_frame* out) {
  const DevBuf& f = s.pool[buf];
  out->width = s.W; out->height = s.H; out->cwidth = s.CW; out->cheight = s.CH;
  out->timestamp_ns = ts;
  out->ring_slot = slot;
  if (on_device) {
    out->y = f.p.y; out->cb = f.p.cb; out->cr = f.p.cr;
    out->y_stride = f.p.pw; out->c_stride = f.p.cw;
    out->on_device = 1;
  } else {
    uint8_t* h = s.host_ring[slot];
    out->y = h;
    out->cb = h + static_cast<size_t>(s.ystride) * s.H;
    out->cr = out->cb + static_cast<size_t>(s.cstride) * s.CH;
    out->y_stride = s.ystride; out->c_stride = s.cstride;
    out->on_device = 0;
  }
}

// storeImage into the host ring (frame.go:67-77): visible region only.
int download(Engine& e, Session& s, int buf, int slot) {
  if (int rc = s.ensure_host_ring()) return rc;
  const DevBuf& f = s.pool[buf];
  uint8_t* h = s.host_ring[slot];
  uint8_t* hcb = h + static_cast<size_t>(s.ystride) * s.H;
  uint8_t* hcr = hcb + static_cast<size_t>(s.cstride) * s.CH;
  if (hipMemcpy2DAsync(h, s.ystride, f.p.y, f.p.pw, s.W, s.H, hipMemcpyDeviceToHost, e.stream) != hipSuccess ||
      hipMemcpy2DAsync(hcb, s.cstride, f.p.cb, f.p.cw, s.CW, s.CH, hipMemcpyDeviceToHost, e.stream) != hipSuccess ||
      hipMemcpy2DAsync(hcr, s.cstride, f.p.cr, f.p.cw, s.CW, s.CH, hipMemcpyDeviceToHost, e.stream) != hipSuccess ||
      hipStreamSynchronize(e.stream) != hipSuccess)
    return s.fail(EBIV_ERR_DEVICE, "frame download failed");
  return EBIV_OK;
}

int engine_init(Engine& e, int device) {
  if (int rc = ensure_device(device)) return rc;
  e.device = device;
  if (hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking) != hipSuccess) {
    e.stream = nullptr;
    return EBIV_ERR_DEVICE;
  }
  if (hipStreamCreateWithFlags(&e.pstream, hipStreamNonBlocking) != hipSuccess) {
    e.pstream = nullptr;
    return EBIV_ERR_DEVICE;
  }
  return EBIV_OK;
}

// Codec.Decode (codec.go:70-111) as a one-frame step.
int decode_one(ebiv_codec* c, const uint8_t* data, const uint8_t* dev, size_t len, int64_t ts, ebiv_frame* out) {
  Session& s = *c->s;
  Engine& e = c->eng;
  if (!data && len) return s.fail(EBIV_ERR_ARGUMENT, "null packet");
  if (hipSetDevice(e.device) != hipSuccess) return s.fail(EBIV_ERR_DEVICE, "hipSetDevice failed");
  StepBuf& S = e.steps[0];
  S.clear();
  S.ss.push_back(&s);
  S.frames.resize(1);
  S.frames[0].emplace_back();
  StepFrame& f = S.frames[0][0];
  f.host = data;
  f.dev = dev;
  f.len = len;
  f.ts = ts;
  int next_slot = 0;
  s.prepare_frame(f, &next_slot);
  if (f.rc) return s.fail(f.rc, f.msg);
  const int rc = e.run_plain(S);
  S.live = false;
  if (rc) return s.fail(rc, e.msg);
  if (f.rc) return s.fail(f.rc, f.msg);
  if (!dev) {
    if (int rc = download(e, s, f.out_buf, f.ring_slot)) return rc;
  }
  fill_frame(s, f.out_buf, f.ring_slot, ts, dev != nullptr, out);
  return EBIV_OK;
}

}  // namespace

extern "C" {

int ebiv_codec_create(int ring_size, int device, ebiv_codec** out) {
  if (!out) return EBIV_ERR_ARGUMENT;
  *out = nullptr;
  auto* c = new ebiv_codec();
  if (int rc = engine_init(c->eng, device)) {
    delete c;
    return rc;
  }
  c->eng.sessions.emplace_back(new Session());
  c->s = c->eng.sessions.back().get();
  c->s->ring_size = ring_size >= 2 ? ring_size : 3;
  c->s->window = 1;
  c->s->stream = c->eng.stream;
  *out = c;
  return EBIV_OK;
}

int ebiv_codec_decode(ebiv_codec* c, const uint8_t* data, size_t len, int64_t ts, ebiv_frame* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  return decode_one(c, data, nullptr, len, ts, out);
}

int ebiv_codec_decode_device(ebiv_codec* c, const uint8_t* host_data, const uint8_t* dev_data, size_t len, int64_t ts,
                             ebiv_frame* out) {
  if (!c || !out || (!dev_data && len)) return EBIV_ERR_ARGUMENT;
  return decode_one(c, host_data, dev_data, len, ts, out);
}

void ebiv_codec_flush(ebiv_codec* c) {
  if (c) c->s->next = 0;
}

void ebiv_codec_destroy(ebiv_codec* c) { delete c; }

const char* ebiv_codec_last_error(const ebiv_codec* c) { return c ? c->s->msg.c_str() : "null codec"; }

int ebiv_codec_stats(const ebiv_codec* c, ebiv_stats* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  *out = c->s->stats;
  return EBIV_OK;
}

int ebiv_codec_debug_dump(ebiv_codec* c, int slot, void* recs, size_t recs_bytes, uint32_t* coefs, size_t coef_count,
                          void* status, size_t status_bytes) {
  if (!c || slot < 0 || slot > 1 || !c->s->ready || static_cast<int>(c->s->slots.size()) <= slot)
    return -EBIV_ERR_ARGUMENT;
  Session& s = *c->s;
  if (hipSetDevice(c->eng.device) != hipSuccess) return -EBIV_ERR_DEVICE;
  const PayloadSlot* ps = s.slots[slot].get();
  const size_t nmb = static_cast<size_t>(s.mbc) * s.mbr;
  if (recs && ps->recs.d &&
      hipMemcpy(recs, ps->recs.d, std::min(recs_bytes, nmb * sizeof(MBRec)), hipMemcpyDeviceToHost) != hipSuccess)
    return -EBIV_ERR_DEVICE;
  if (coefs && ps->coefs.d &&
      hipMemcpy(coefs, ps->coefs.d, std::min(coef_count, nmb * kCoefPerMB) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return -EBIV_ERR_DEVICE;
  if (status && ps->h_status.h)
    std::memcpy(status, ps->h_status.h, std::min(status_bytes, ps->h_status.cap * sizeof(TileStatus)));
  return static_cast<int>(nmb);
}

int ebiv_codec_set_timing(ebiv_codec* c, int enabled) {
  if (!c) return EBIV_ERR_ARGUMENT;
  c->eng.timing = enabled != 0;
  c->eng.ktimes = ebiv_kernel_times{};
  return EBIV_OK;
}

int ebiv_codec_kernel_times(const ebiv_codec* c, ebiv_kernel_times* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  *out = c->eng.ktimes;
  return EBIV_OK;
}

// The device buffer behind a frame this engine returned: a device frame by
// its plane pointer, a host frame by the ring slot whose image it views.
static const DevBuf* buffer_of(const Engine& e, const ebiv_frame* f) {
  for (auto& s : e.sessions) {
    if (!s->ready) continue;
    if (f->on_device) {
      for (auto& p : s->pool)
        if (p.p.y == f->y) return &p;
    } else if (f->ring_slot >= 0 && f->ring_slot < s->ring_size && !s->host_ring.empty() &&
               s->host_ring[f->ring_slot] == f->y && s->ring[f->ring_slot] >= 0) {
      return &s->pool[s->ring[f->ring_slot]];
    }
  }
  return nullptr;
}

static RgbaJob rgba_job(const DevBuf& b, int w, int h, uint8_t* dst, int64_t pitch) {
  return RgbaJob{b.p.y, b.p.cb, b.p.cr, dst, pitch, b.p.pw, b.p.cw, w, h};
}

static int frame_rgba(Engine& e, const ebiv_frame* f, uint8_t* dst, int dst_on_device) {
  if (!f || !dst || !f->y || f->width <= 0 || f->height <= 0) return EBIV_ERR_ARGUMENT;
  if (hipSetDevice(e.device) != hipSuccess) return EBIV_ERR_DEVICE;
  const DevBuf* b = buffer_of(e, f);
  if (!b) return EBIV_ERR_ARGUMENT;
  const size_t px = static_cast<size_t>(f->width) * f->height;
  uint32_t* out = reinterpret_cast<uint32_t*>(dst);
  if (!dst_on_device) {
    if (!e.d_rgba.reserve(px)) return EBIV_ERR_DEVICE;
    out = e.d_rgba.d;
  }
  const int m = e.mark_begin(3);
  launch_rgba(e.stream, rgba_job(*b, f->width, f->height, reinterpret_cast<uint8_t*>(out), 4 * static_cast<int64_t>(f->width)));
  e.mark_end(m);
  if (hipGetLastError() != hipSuccess) return EBIV_ERR_DEVICE;
  if (!dst_on_device && hipMemcpyAsync(dst, e.d_rgba.d, px * 4, hipMemcpyDeviceToHost, e.stream) != hipSuccess)
    return EBIV_ERR_DEVICE;
  if (hipStreamSynchronize(e.stream) != hipSuccess) return EBIV_ERR_DEVICE;
  e.collect_marks();
  return EBIV_OK;
}

int ebiv_frame_rgba(ebiv_codec* c, const ebiv_frame* f, uint8_t* dst, int dst_on_device) {
  if (!c) return EBIV_ERR_ARGUMENT;
  return frame_rgba(c->eng, f, dst, dst_on_device);
}

int ebiv_frame_rgba_noalloc(ebiv_codec* c, const ebiv_frame* f, uint8_t* dst) {
  if (!c) return EBIV_ERR_ARGUMENT;
  Session& s = *c->s;
  Engine& e = c->eng;
  if (!f || !dst || !f->y || f->width <= 0 || f->height <= 0) return s.fail(EBIV_ERR_ARGUMENT, "null frame or destination");
  const size_t bytes = static_cast<size_t>(f->width) * static_cast<size_t>(f->height) * 4;
  if (bytes > kRgbaScratchBytes)
    return s.fail(EBIV_ERR_ARGUMENT, "frame too large for ebiv_frame_rgba_noalloc (at most 3840x2160 pixels)");
  if (hipSetDevice(e.device) != hipSuccess) return s.fail(EBIV_ERR_DEVICE, "hipSetDevice failed");
  uint8_t* stage = e.device < static_cast<int>(g_scratch.size()) ? g_scratch[e.device] : nullptr;
  if (!stage) return s.fail(EBIV_ERR_DEVICE, "device not initialised");
  const DevBuf* b = buffer_of(e, f);
  if (!b) return s.fail(EBIV_ERR_ARGUMENT, "frame not returned by this codec");
  // The kernel writes all `bytes` of the staging area that are copied out,
  // so nothing needs clearing between calls.
  launch_rgba(e.stream, rgba_job(*b, f->width, f->height, stage, 4 * static_cast<int64_t>(f->width)));
  if (hipGetLastError() != hipSuccess) return s.fail(EBIV_ERR_DEVICE, "kernel launch failed");
  if (hipMemcpyAsync(dst, stage, bytes, hipMemcpyDeviceToHost, e.stream) != hipSuccess ||
      hipStreamSynchronize(e.stream) != hipSuccess)
    return s.fail(EBIV_ERR_DEVICE, "rgba download failed");
  return EBIV_OK;
}

// ---- batched engine ------------------------------------------------------------
int ebiv_batch_create(int device, ebiv_batch** out) {
  if (!out) return EBIV_ERR_ARGUMENT;
  *out = nullptr;
  auto* b = new ebiv_batch();
  if (int rc = engine_init(b->eng, device)) {
    delete b;
    return rc;
  }
  *out = b;
  return EBIV_OK;
}

void ebiv_batch_destroy(ebiv_batch* b) { delete b; }

int ebiv_batch_open(ebiv_batch* b, int ring_size, int window, int* session) {
  if (!b || !session || window < 1 || window > 64) return EBIV_ERR_ARGUMENT;
  auto s = std::make_unique<Session>();
  s->id = b->eng.next_id++;
  s->window = window;
  // every frame a step returns stays valid until the next step returns
  s->ring_size = std::max(ring_size >= 2 ? ring_size : 3, window);
  s->stream = b->eng.stream;
  *session = s->id;
  b->eng.sessions.push_back(std::move(s));
  return EBIV_OK;
}

int ebiv_batch_close(ebiv_batch* b, int session) {
  if (!b) return EBIV_ERR_ARGUMENT;
  Engine& e = b->eng;
  auto& v = e.sessions;
  for (size_t i = 0; i < v.size(); i++)
    if (v[i]->id == session) {
      (void)hipSetDevice(e.device);
      (void)hipStreamSynchronize(e.pstream);
      (void)hipStreamSynchronize(e.stream);
      // A step in flight (its entropy launch queued by the last call) drops
      // the session's frames: no completions for a closed session.
      for (auto& S : e.steps)
        for (auto& p : S.ss)
          if (p == v[i].get()) p = nullptr;
      e.carry.erase(std::remove_if(e.carry.begin(), e.carry.end(),
                                   [&](const ebiv_completion& c) { return c.session == session; }),
                    e.carry.end());
      v.erase(v.begin() + static_cast<long>(i));
      return EBIV_OK;
    }
  return EBIV_ERR_ARGUMENT;
}

int ebiv_batch_submit(ebiv_batch* b, int session, const uint8_t* host_data, const uint8_t* dev_data, size_t len,
                      int64_t ts, uint64_t tag) {
  if (!b || (!host_data && len)) return EBIV_ERR_ARGUMENT;
  Session* s = b->eng.find(session);
  if (!s) return EBIV_ERR_ARGUMENT;
  Session::Queued q;
  q.host.assign(host_data, host_data + len);
  q.dev = dev_data;
  q.ts = ts;
  q.tag = tag;
  s->queue.push_back(std::move(q));
  return EBIV_OK;
}

int ebiv_batch_submit_many(ebiv_batch* b, int session, int n, const uint8_t* const* host_data,
                           const uint8_t* const* dev_data, const size_t* len, const int64_t* ts, const uint64_t* tags) {
  if (!b || n < 0 || (n > 0 && (!host_data || !len))) return EBIV_ERR_ARGUMENT;
  if (!b->eng.find(session)) return EBIV_ERR_ARGUMENT;
  for (int i = 0; i < n; i++)
    if (!host_data[i] && len[i]) return EBIV_ERR_ARGUMENT;  // (checked before anything is queued)
  for (int i = 0; i < n; i++) {
    const int rc = ebiv_batch_submit(b, session, host_data[i], dev_data ? dev_data[i] : nullptr, len[i],
                                     ts ? ts[i] : 0, tags ? tags[i] : 0);
    if (rc) return rc;
  }
  return EBIV_OK;
}

int ebiv_batch_flush(ebiv_batch* b, int session) {
  Session* s = b ? b->eng.find(session) : nullptr;
  if (!s) return EBIV_ERR_ARGUMENT;
  s->queue.clear();
  s->next = 0;  // Codec.Flush
  return EBIV_OK;
}

int ebiv_batch_pending(const ebiv_batch* b, int session) {
  if (!b) return 0;
  const Engine& e = b->eng;
  int n = 0;
  for (auto& s : e.sessions)
    if (session == 0 || s->id == session) n += static_cast<int>(s->queue.size());
  for (const auto& S : e.steps)  // taken by a step in flight
    if (S.live) n += static_cast<int>(S.frames_of(session));
  for (const auto& c : e.carry)  // decoded, not yet returned
    if (session == 0 || c.session == session) n++;
  return n;
}

namespace {

constexpr int kSlotsPerStep = 128;  // payload slots of a session per step buffer (window + an override)

// Takes up to a window of queued packets from every session into S (at most
// `budget` in all), starting one session further on each step so that a
// budget below the queued work is shared, and parses their headers and
// tables, sessions in parallel, each in stream order.  A key frame with a
// golden override ends its session's window: the packets after it go back
// to the front of the queue.
void take_step(Engine& e, StepBuf& S, int budget, int slot_base) {
  S.clear();
  const size_t ns = e.sessions.size();
  const size_t first = ns ? e.rotate++ % ns : 0;
  for (size_t n = 0; n < ns; n++) {
    Session& s = *e.sessions[(first + n) % ns];
    if (s.queue.empty() || budget <= 0) continue;
    S.ss.push_back(&s);
    S.frames.emplace_back();
    S.held.emplace_back();
    const int window = std::min(s.window, budget);
    S.frames.back().reserve(static_cast<size_t>(window));
    S.held.back().reserve(static_cast<size_t>(window));
    while (!s.queue.empty() && static_cast<int>(S.held.back().size()) < window) {
      S.held.back().push_back(std::move(s.queue.front()));
      s.queue.pop_front();
      budget--;
    }
  }
  if (S.ss.empty()) return;
  if (!e.pool) e.pool.reset(new HostPool(e.device));
  e.pool->parallel_for(static_cast<int>(S.ss.size()), [&](int i) {
    Session& s = *S.ss[static_cast<size_t>(i)];
    std::vector<StepFrame>& fr = S.frames[static_cast<size_t>(i)];
    std::vector<Session::Queued>& h = S.held[static_cast<size_t>(i)];
    int next_slot = slot_base;
    size_t k = 0;
    while (k < h.size()) {
      const Session::Queued& q = h[k++];
      fr.emplace_back();
      StepFrame& f = fr.back();
      f.host = q.host.data();
      f.dev = q.dev;
      f.len = q.host.size();
      f.ts = q.ts;
      f.tag = q.tag;
      if (!s.prepare_frame(f, &next_slot)) break;
    }
    for (size_t j = h.size(); j > k; j--) s.queue.push_front(std::move(h[j - 1]));
    h.resize(k);
  });
}

}  // namespace

// A step: the entropy launch of the current step buffer (already in flight
// when the previous call launched it), its tile statuses, reconstruction
// from the real outcomes and Go's state transitions; then, with the
// pipeline on, the next step's packets are taken and its entropy launch
// queued on the entropy stream, where it runs beside this step's
// reconstruction; then this step's completions.
int ebiv_batch_step(ebiv_batch* b, ebiv_completion* out, int max_out, int* n_out) {
  if (!b || !n_out || max_out < 0 || (max_out > 0 && !out)) return EBIV_ERR_ARGUMENT;
  *n_out = 0;
  Engine& e = b->eng;
  if (hipSetDevice(e.device) != hipSuccess) return e.fail(EBIV_ERR_DEVICE, "hipSetDevice failed");
  if (!e.carry.empty()) {  // completions an earlier call had no room for
    const size_t give = std::min(e.carry.size(), static_cast<size_t>(max_out));
    std::copy(e.carry.begin(), e.carry.begin() + static_cast<long>(give), out);
    e.carry.erase(e.carry.begin(), e.carry.begin() + static_cast<long>(give));
    *n_out = static_cast<int>(give);
    return EBIV_OK;
  }
  const bool pipe = e.sched.pipeline;
  hipStream_t ps = pipe ? e.pstream : e.stream;
  StepBuf& S = e.steps[e.cur];
  if (!S.live) {
    take_step(e, S, max_out, e.cur * kSlotsPerStep);
    if (S.ss.empty()) return EBIV_OK;
    if (int rc = e.launch_entropy(S, ps, false)) return rc;
  }
  if (int rc = e.finish_entropy(S)) return rc;
  if (int rc = e.launch_recon(S)) return rc;
  e.commit(S);
  if (pipe) {
    const int nx = e.cur ^ 1;
    StepBuf& N = e.steps[nx];
    take_step(e, N, max_out, nx * kSlotsPerStep);
    // Only a step that launched moves the engine to the other buffer set: a
    // lone step (every queued frame fits the window) keeps reusing one set,
    // whose device buffers are already sized, instead of allocating both.
    if (!N.ss.empty()) {
      if (int rc = e.launch_entropy(N, ps, false)) return rc;
      e.cur = nx;
    }
  }
  if (int rc = e.wait_recon()) return rc;
  size_t npresent = 0, nframes = 0;
  uint32_t max_px8 = 0;
  for (auto& fr : S.frames) nframes += fr.size();
  // (reserve does not keep contents: once, before the table is filled)
  if (!e.h_present.reserve(nframes + 1)) return e.fail(EBIV_ERR_DEVICE, "presentation table allocation failed");
  std::vector<ebiv_completion> done;
  done.reserve(nframes);
  for (size_t i = 0; i < S.ss.size(); i++) {
    if (!S.ss[i]) continue;  // closed while in flight
    for (auto& f : S.frames[i]) {
      done.emplace_back();
      ebiv_completion& c = done.back();
      std::memset(&c, 0, sizeof c);
      c.session = S.ss[i]->id;
      c.status = f.rc;
      c.tag = f.tag;
      c.surface = -1;
      if (f.rc == EBIV_OK) {
        Session& s = *S.ss[i];
        fill_frame(s, f.out_buf, f.ring_slot, f.ts, true, &c.frame);
        if (!s.surfaces.empty() && s.surface_pitch >= 4 * static_cast<size_t>(s.W)) {
          c.surface = s.next_surface;
          s.next_surface = (s.next_surface + 1) % static_cast<int>(s.surfaces.size());
          e.h_present.h[npresent++] = rgba_job(s.pool[static_cast<size_t>(f.out_buf)], s.W, s.H,
                                               s.surfaces[static_cast<size_t>(c.surface)],
                                               static_cast<int64_t>(s.surface_pitch));
          max_px8 = std::max(max_px8, static_cast<uint32_t>((s.W + 7) / 8) * static_cast<uint32_t>(s.H));
        }
      } else {
        std::strncpy(c.message, f.msg.c_str(), sizeof c.message - 1);
        S.ss[i]->msg = f.msg;
      }
    }
  }
  // ---- presentation: every presented frame of the step in one launch
  if (npresent) {
    if (!e.d_present.reserve(npresent)) return e.fail(EBIV_ERR_DEVICE, "presentation table allocation failed");
    if (hipMemcpyAsync(e.d_present.d, e.h_present.h, sizeof(RgbaJob) * npresent, hipMemcpyHostToDevice, e.stream) !=
        hipSuccess)
      return e.fail(EBIV_ERR_DEVICE, "presentation table upload failed");
    const int m = e.mark_begin(3);
    launch_rgba_frames(e.stream, e.d_present.d, static_cast<int>(npresent), max_px8);
    e.mark_end(m);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e.stream) != hipSuccess)
      return e.fail(EBIV_ERR_DEVICE, "presentation failed on the device");
    e.collect_marks();
  }
  S.clear();
  const size_t give = std::min(done.size(), static_cast<size_t>(max_out));
  std::copy(done.begin(), done.begin() + static_cast<long>(give), out);
  e.carry.insert(e.carry.end(), done.begin() + static_cast<long>(give), done.end());
  *n_out = static_cast<int>(give);
  return EBIV_OK;
}


const char* ebiv_batch_last_error(const ebiv_batch* b) { return b ? b->eng.msg.c_str() : "null batch"; }

int ebiv_batch_set_timing(ebiv_batch* b, int enabled) {
  if (!b) return EBIV_ERR_ARGUMENT;
  b->eng.timing = enabled != 0;
  b->eng.ktimes = ebiv_kernel_times{};
  return EBIV_OK;
}

int ebiv_batch_kernel_times(const ebiv_batch* b, ebiv_kernel_times* out) {
  if (!b || !out) return EBIV_ERR_ARGUMENT;
  *out = b->eng.ktimes;
  return EBIV_OK;
}

int ebiv_batch_stats(const ebiv_batch* b, int session, ebiv_stats* out) {
  if (!b || !out) return EBIV_ERR_ARGUMENT;
  ebiv_stats t{0, 0, 0, 0};
  for (auto& s : b->eng.sessions)
    if (session == 0 || s->id == session) {
      t.frames += s->stats.frames;
      t.symbols += s->stats.symbols;
      t.tiles += s->stats.tiles;
      t.tiles_generic += s->stats.tiles_generic;
    }
  *out = t;
  return EBIV_OK;
}

int ebiv_batch_set_surfaces(ebiv_batch* b, int session, uint8_t* const* surfaces, int count, size_t pitch) {
  Session* s = b ? b->eng.find(session) : nullptr;
  // fewer surfaces than the window would let one step overwrite its own frames
  if (!s || count < 0 || count > 64 || (count > 0 && (count < s->window || !surfaces || pitch == 0)))
    return EBIV_ERR_ARGUMENT;
  for (int i = 0; i < count; i++)
    if (!surfaces[i]) return EBIV_ERR_ARGUMENT;
  s->surfaces.assign(surfaces, surfaces + count);
  s->surface_pitch = pitch;
  s->next_surface = 0;
  return EBIV_OK;
}

int ebiv_batch_set_schedule(ebiv_batch* b, const ebiv_schedule* sc) {
  if (!b || !sc) return EBIV_ERR_ARGUMENT;
  Schedule& d = b->eng.sched;
  auto waves = [](int32_t v, int& dst) { if (v > 0) dst = v > 16 ? 16 : v; };
  waves(sc->intra_waves_key, d.intra_waves_key);
  waves(sc->intra_waves_key_many, d.intra_waves_key_many);
  waves(sc->intra_waves_few, d.intra_waves_few);
  waves(sc->intra_waves_many, d.intra_waves_many);
  if (sc->many_tiles > 0) d.many_tiles = sc->many_tiles;
  if (sc->recon_lanes > 0) d.recon_lanes = sc->recon_lanes > Engine::kMaxLanes ? Engine::kMaxLanes : sc->recon_lanes;
  // (a step in flight is finished by the next call either way: it waits for
  // its entropy launch's event, on whichever stream that ran)
  if (sc->pipeline > 2) return EBIV_ERR_ARGUMENT;
  if (sc->pipeline > 0) d.pipeline = sc->pipeline == 1;
  if (sc->intra_cells > 2) return EBIV_ERR_ARGUMENT;
  if (sc->intra_cells > 0) d.intra_cells = sc->intra_cells == 1;
  if (sc->key_band_rows > 0) d.key_band_rows = sc->key_band_rows > kFrontMax ? kFrontMax : sc->key_band_rows;
  return EBIV_OK;
}

int ebiv_batch_get_schedule(const ebiv_batch* b, ebiv_schedule* out) {
  if (!b || !out) return EBIV_ERR_ARGUMENT;
  const Schedule& d = b->eng.sched;
  *out = ebiv_schedule{d.intra_waves_key, d.intra_waves_few, d.intra_waves_many, d.many_tiles, d.recon_lanes,
                       d.pipeline ? 1 : 2, d.intra_waves_key_many, d.intra_cells ? 1 : 2,
                       d.key_band_rows};
  return EBIV_OK;
}

int ebiv_batch_frame_rgba(ebiv_batch* b, const ebiv_frame* f, uint8_t* dst, int dst_on_device) {
  if (!b) return EBIV_ERR_ARGUMENT;
  return frame_rgba(b->eng, f, dst, dst_on_device);
}

size_t ebiv_session_bytes(int width, int height, int ring_size, int window) {
  if (width <= 0 || height <= 0) return 0;
  const size_t mbc = static_cast<size_t>((width + kMB - 1) / kMB), mbr = static_cast<size_t>((height + kMB - 1) / kMB);
  const size_t ysz = mbc * kMB * mbr * kMB, csz = align256(mbc * kCMB * mbr * kCMB);
  const size_t rs = static_cast<size_t>(std::max(ring_size >= 2 ? ring_size : 3, window));
  const size_t pool = (rs + 2 + static_cast<size_t>(window) + 1) * (ysz + 2 * csz);
  // payload slots: a window's payloads plus an override, in each of the
  // engine's two step buffers (one reconstructing while the next parses)
  const size_t payload =
      2 * (static_cast<size_t>(window) + 1) * mbc * mbr * (sizeof(MBRec) + kCoefPerMB * 4 + (kResPerMB + kResSyncPerMB) * 2);
  return pool + payload;
}

int ebiv_device_alloc(int device, size_t bytes, void** out) {
  if (!out) return EBIV_ERR_ARGUMENT;
  if (int rc = ensure_device(device)) return rc;
  return hipMalloc(out, bytes + 256) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_device_free(void* p) { return hipFree(p) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE; }
int ebiv_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"

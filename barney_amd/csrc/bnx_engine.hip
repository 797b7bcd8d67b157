// bnx_engine.hip -- host side of the wavefront: the per-request frame loop
// (bnx_render) and the request server (bnx_server_*).  Both drive the same
// generation: trace, volume march and shade of one shared ray queue, whose
// rays may belong to any number of requests.
//
// Queue layout: the queue of a generation is a list of segments, each a run
// of rays of one request (Segment, bnx_render.h).  Shade writes the rays of
// request s into s's own region of the other queue, sized for the worst case
// (each path leaves at most one shadow ray and one continuation ray), and
// counts them per request; the next generation's segments are those regions
// plus the runs of new primary rays that top the queue up.  Every block of
// every launch lies inside one segment, so a request's parameters are
// block-uniform and live in scalar registers.
//
// Replaces Context::renderTiles (native/Context.cpp:139-182) and the
// RayQueue count read-back (render/RayQueue.cpp:47-57).  The server has no
// counterpart in the reference: it continuously batches requests of
// different worlds, resolutions and spp through that one wavefront.
#include "bnx_internal.h"
#include "bnx_render.h"
#include "bnx_volume.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace bnx {
namespace {

// pinned host memory for read-backs and small uploads, or (borrowed) a
// fixed slice of host memory owned elsewhere
struct Pinned {
  void *p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  ~Pinned()
  {
    if (p && !borrowed) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t b)
  {
    if (borrowed) return b <= bytes ? hipSuccess : hipErrorOutOfMemory;
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, b ? b : 16, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void borrow(void *q, size_t b)
  {
    p = q;
    bytes = b;
    borrowed = true;
  }
  template <typename T> T *as() const { return (T *)p; }
};

// device-driven generations of one batch at most (a captured bounce loop)
constexpr int MAX_DD_GENS = 16;
// smallest generation worth reordering
constexpr int SORT_MIN_RAYS = 1 << 16;

// what the launches of a device-driven batch depend on, besides the
// wavefront's buffers: every tuning field that selects a kernel variant or a
// launch size inside the captured loop (a graph replays the kernels it was
// captured with)
struct BatchShape {
  long long count0 = 0, cap = 0;
  bool surfaces = false, volumes = false;
  int full = 0;       // the march kernel: 0 lean, else MARCH_FULL / MARCH_CELLS bits
  int refill = 64, stack = 16, walk = 0, blocks_per_cu = 8, shade_block = 1024;
  bool ext = false;   // the trace kernels with the cylinder test and the cut plane
  bool xf = false;    // the trace kernels that change the ray's frame at instances
  bool operator<(const BatchShape &o) const
  {
    return std::tie(count0, cap, surfaces, volumes, full, refill, stack, walk, blocks_per_cu, shade_block, ext, xf) <
           std::tie(o.count0, o.cap, o.surfaces, o.volumes, o.full, o.refill, o.stack, o.walk, o.blocks_per_cu,
                    o.shade_block, o.ext, o.xf);
  }
};

// a captured bounce loop: G generations of one batch shape
struct BatchGraph {
  BatchShape shape;
  int G = 0, cur0 = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// queues, tables and staging of one wavefront
struct Wavefront {
  Queues Q;
  DevMem slots, chunks, segs, blk256, blk1024, blkT, pool, out_base, out_cnt;
  Pinned h_slots, h_chunks, h_segs, h_out_base, h_out_cnt;
  int slot_cap = 0, chunk_cap = 0, seg_cap = 0;
  // device-driven bounce loops: per-generation counts, timing events (four
  // per generation), generations observed per batch shape, captured graphs
  // (valid for the buffer addresses they were captured with)
  DevMem gen_counts, stamps;
  Pinned h_gen_counts;
  SortScratch sorter;
  std::map<BatchShape, int> gens_seen;
  std::vector<BatchGraph> graphs;
  std::vector<const void *> graph_bufs;
  bool graphs_off = false;      // capture failed once: enqueue directly
  hipError_t capture_err = hipSuccess;   // why (reported in every render's stats)
  // every buffer is a fixed slice of the static scratch array (device) and
  // of host_stage (host): nothing is allocated, no timing events exist, and
  // the paths per batch are bounded by the slices
  bool noalloc = false;
  // host copy of the statistics counters (read back once per render)
  unsigned long long h_stats[STAT_COUNT * STAT_SLOTS];
  ~Wavefront() { drop_graphs(); }
  void drop_graphs()
  {
    for (BatchGraph &g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs.clear();
  }
  // the addresses captured graphs refer to; a reallocation drops the graphs
  void check_graph_bufs()
  {
    std::vector<const void *> b = {Q.org[0].ptr, Q.org[1].ptr, Q.dir[0].ptr, Q.dir[1].ptr, Q.rng[0].ptr,
                                   Q.rng[1].ptr, Q.ps[0].ptr, Q.ps[1].ptr, Q.hit[0].ptr, Q.hit[1].ptr,
                                   Q.hit_v[0].ptr, Q.hit_v[1].ptr, Q.vcol[0].ptr, Q.vcol[1].ptr, Q.stats.ptr,
                                   slots.ptr, segs.ptr, blk256.ptr, blk1024.ptr, blkT.ptr, pool.ptr,
                                   out_base.ptr, out_cnt.ptr, gen_counts.ptr, stamps.ptr};
    if (b != graph_bufs) {
      drop_graphs();
      graph_bufs = b;
    }
  }
  // host staging: G counts, 4 G stamps, then the two output counters
  int ensure_dd()
  {
    BNX_HIP(gen_counts.ensure(sizeof(long long) * MAX_DD_GENS));
    BNX_HIP(stamps.ensure(sizeof(unsigned long long) * 4 * MAX_DD_GENS));
    BNX_HIP(h_gen_counts.ensure(sizeof(long long) * 5 * MAX_DD_GENS + sizeof(int) * 2));
    return BNX_OK;
  }
  size_t blk_cap = 0;

  // rays per queue, request slots, raygen chunks, segments per generation
  int ensure(size_t rays, int nslots, int nchunks, int nsegs)
  {
    int rc = Q.ensure(rays);
    if (rc != BNX_OK) return rc;
    if (!Q.ev_begin && !noalloc) {
      BNX_HIP(hipEventCreate(&Q.ev_begin));
      BNX_HIP(hipEventCreate(&Q.ev_end));
      BNX_HIP(hipEventCreate(&Q.ev_vbegin));
      BNX_HIP(hipEventCreate(&Q.ev_vend));
    }
    if (nslots > slot_cap) {
      BNX_HIP(slots.ensure(sizeof(SlotDev) * nslots));
      BNX_HIP(h_slots.ensure(sizeof(SlotDev) * nslots));
      BNX_HIP(out_base.ensure(sizeof(int) * nslots));
      BNX_HIP(h_out_base.ensure(sizeof(int) * nslots));
      BNX_HIP(out_cnt.ensure(sizeof(int) * 2 * nslots));
      BNX_HIP(h_out_cnt.ensure(sizeof(int) * 2 * nslots));
      slot_cap = nslots;
    }
    if (nchunks > chunk_cap) {
      BNX_HIP(chunks.ensure(sizeof(RaygenChunk) * nchunks));
      BNX_HIP(h_chunks.ensure(sizeof(RaygenChunk) * nchunks));
      chunk_cap = nchunks;
    }
    if (nsegs > seg_cap) {
      BNX_HIP(segs.ensure(sizeof(Segment) * nsegs));
      BNX_HIP(h_segs.ensure(sizeof(Segment) * nsegs));
      BNX_HIP(pool.ensure(sizeof(int) * nsegs));
      seg_cap = nsegs;
    }
    // every segment adds at most one partial block
    const size_t blocks = rays / 256 + (size_t)nsegs + 1;
    if (blocks > blk_cap) {
      BNX_HIP(blk256.ensure(sizeof(int) * blocks));
      BNX_HIP(blk1024.ensure(sizeof(int) * blocks));
      BNX_HIP(blkT.ensure(sizeof(int) * blocks));
      blk_cap = blocks;
    }
    return BNX_OK;
  }
};

// segments of one generation, written straight into pinned staging
struct SegList {
  Segment *v;
  int n = 0, blocks256 = 0, blocks1024 = 0, blocksT = 0;
  long long rays = 0;
  // refill: idle lanes per wave that trigger a trace-pool refill.  Whole-
  // wave refills (64) keep consecutive primary rays together; secondary
  // rays are incoherent anyway, and refilling lanes as they finish keeps
  // the waves busy
  void add(int slot, int begin, int count, int refill)
  {
    if (count <= 0) return;
    v[n++] = Segment{slot, begin, count, blocks256, blocks1024, 0, refill, 0};
    blocks256 += (count + 255) / 256;
    blocks1024 += (count + 1023) / 1024;
    rays += count;
  }
  // trace blocks: each draws rays from its segment's pool; size the launch
  // to about `target` blocks (a few per CU), at least one ray per thread
  void layout_trace(int target)
  {
    const long long per = 256LL * std::max(1LL, (rays + 256LL * target - 1) / (256LL * target));
    blocksT = 0;
    for (int k = 0; k < n; k++) {
      v[k].blockT = blocksT;
      blocksT += (int)((v[k].count + per - 1) / per);
    }
  }
};

struct GenTiming {
  double trace_ms = 0.0, march_ms = 0.0;
  int64_t trace_launches = 0, march_launches = 0, march_rays = 0;
};

// Enqueue one generation over queue `cur`: upload the segments, build the
// block maps, trace, volume march, shade into queue 1 - cur.  The caller
// has uploaded the output region bases and cleared the output counters.
int enqueue_generation(Device *dev, Wavefront &F, int cur, SegList &sl, bool surfaces, bool volumes, int full)
{
  Queues &Q = F.Q;
  RayQueue in = Q.q(cur), out = Q.q(1 - cur);
  unsigned long long *st = Q.stats.as<unsigned long long>();
  const SlotDev *slots = F.slots.as<SlotDev>();
  int rc;
  sl.layout_trace(dev->trace_blocks_per_cu * dev->num_cus);
  BNX_HIP(hipMemcpyAsync(F.segs.ptr, sl.v, sizeof(Segment) * sl.n, hipMemcpyHostToDevice, dev->stream));
  const SegView sv{F.segs.as<Segment>(), F.blk256.as<int>(), F.blk1024.as<int>(), F.blkT.as<int>(),
                   F.pool.as<int>()};
  if ((rc = launch_segment_blocks(dev, sv, sl.n, sl.blocks256, sl.blocks1024, sl.blocksT)) != BNX_OK) return rc;
  TraceQueue tq{in.org, in.dir, in.hit, in.hit_v, in.rng, in.ps};
  // with neither surfaces nor volumes the trace kernel still writes the misses
  const bool trace = surfaces || !volumes;
  const bool timed = Q.ev_begin != nullptr;
  if (timed) BNX_HIP(hipEventRecord(Q.ev_begin, dev->stream));
  bool pooled = false, cut = false;
  const SlotDev *hs = F.h_slots.as<SlotDev>();
  for (int k = 0; k < sl.n; k++) {
    pooled = pooled || sl.v[k].refill < 64;
    cut = cut || has_cut_plane(hs[sl.v[k].slot].rr);
  }
  if (trace && (rc = launch_trace_queue(dev, slots, sv, sl.blocks256, sl.blocksT, pooled, cut, tq, st)) != BNX_OK)
    return rc;
  if (timed) BNX_HIP(hipEventRecord(Q.ev_end, dev->stream));
  if (volumes) {
    if (timed) BNX_HIP(hipEventRecord(Q.ev_vbegin, dev->stream));
    if ((rc = launch_volume_march(dev, slots, sv, sl.blocks256, in, trace, full, st)) != BNX_OK) return rc;
    if (timed) BNX_HIP(hipEventRecord(Q.ev_vend, dev->stream));
  }
  const SlotOut so{F.out_base.as<int>(), F.out_cnt.as<int>()};
  return launch_shade(dev, slots, sv, sl.blocks256, sl.blocks1024, in, out, so, st);
}

// Enqueue G device-driven generations of one request's batch (slot 0, the
// primary rays in queue cur0, count0 of them): each generation's segment
// and counter reset come from k_gen_setup, and every launch is sized for
// the generation's largest possible count (count0, then cap), the blocks
// past the actual count finding no ray.  No host read-back in between, so
// the sequence can be captured as a graph.  Launch timing: wall-clock
// stamps[4 g ..] before and after the trace and the march launches.
int enqueue_dd(Device *dev, Wavefront &F, int cur0, int G, const BatchShape &b)
{
  Queues &Q = F.Q;
  unsigned long long *st = Q.stats.as<unsigned long long>();
  const SlotDev *slots = F.slots.as<SlotDev>();
  const SegView sv{F.segs.as<Segment>(), F.blk256.as<int>(), F.blk1024.as<int>(), F.blkT.as<int>(),
                   F.pool.as<int>()};
  const SlotOut so{F.out_base.as<int>(), F.out_cnt.as<int>()};
  const bool trace = b.surfaces || !b.volumes;
  const long long target = (long long)dev->trace_blocks_per_cu * dev->num_cus;
  int rc;
  for (int g = 0; g < G; g++) {
    const int cur = (cur0 + g) & 1;
    // primary rays one per thread (whole-wave refills), as in the host loop
    const int refill = g == 0 ? 64 : b.refill;
    const bool pooled = refill < 64;
    const long long c = g == 0 ? b.count0 : b.cap;
    const int blocks256 = (int)((c + 255) / 256), blocks1024 = (int)((c + 1023) / 1024);
    int blocksT = 0;
    if (pooled) {
      const long long per = 256LL * std::max(1LL, (c + 256LL * target - 1) / (256LL * target));
      blocksT = (int)((c + per - 1) / per);
    }
    if ((rc = launch_gen_setup(dev, F.segs.as<Segment>(), F.out_cnt.as<int>(), F.gen_counts.as<long long>(), g,
                               (int)b.count0, refill)) != BNX_OK)
      return rc;
    if ((rc = launch_segment_blocks(dev, sv, 1, blocks256, blocks1024, blocksT)) != BNX_OK) return rc;
    RayQueue in = Q.q(cur), out = Q.q(1 - cur);
    TraceQueue tq{in.org, in.dir, in.hit, in.hit_v, in.rng, in.ps};
    unsigned long long *ts = F.stamps.as<unsigned long long>() + 4 * g;
    if (trace) {
      if ((rc = launch_stamp(dev, ts)) != BNX_OK) return rc;
      if ((rc = launch_trace_queue(dev, slots, sv, blocks256, blocksT, pooled, b.ext, tq, st)) != BNX_OK) return rc;
      if ((rc = launch_stamp(dev, ts + 1)) != BNX_OK) return rc;
    }
    if (b.volumes) {
      if ((rc = launch_stamp(dev, ts + 2)) != BNX_OK) return rc;
      if ((rc = launch_volume_march(dev, slots, sv, blocks256, in, trace, b.full, st)) != BNX_OK) return rc;
      if ((rc = launch_stamp(dev, ts + 3)) != BNX_OK) return rc;
    }
    if ((rc = launch_shade(dev, slots, sv, blocks256, blocks1024, in, out, so, st)) != BNX_OK) return rc;
  }
  return BNX_OK;
}

// the captured bounce loop of shape b with G generations starting in queue
// cur0, captured on first use; nullptr when capture is not available
int batch_graph(Device *dev, Wavefront &F, const BatchShape &b, int G, int cur0, hipGraphExec_t &exec)
{
  exec = nullptr;
  if (F.graphs_off) return BNX_OK;
  F.check_graph_bufs();
  for (BatchGraph &g : F.graphs)
    if (!(g.shape < b) && !(b < g.shape) && g.G == G && g.cur0 == cur0) {
      exec = g.exec;
      return BNX_OK;
    }
  if (F.graphs.size() >= 32) F.drop_graphs();
  BatchGraph bg;
  bg.shape = b;
  bg.G = G;
  bg.cur0 = cur0;
  hipError_t e = hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    F.graphs_off = true;
    F.capture_err = e;
    return BNX_OK;
  }
  const int rc = enqueue_dd(dev, F, cur0, G, b);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(dev->stream, &graph);
  if (rc == BNX_OK && e == hipSuccess && graph) e = hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0);
  if (rc != BNX_OK || e != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (rc != BNX_OK) return rc;            // an enqueue error, not a capture one
    F.graphs_off = true;
    F.capture_err = e != hipSuccess ? e : hipErrorUnknown;
    return BNX_OK;
  }
  bg.graph = graph;
  F.graphs.push_back(bg);
  exec = bg.exec;
  return BNX_OK;
}

// after the stream has synchronized
int collect_timing(Wavefront &F, int n, bool surfaces, bool volumes, GenTiming &t)
{
  float ms = 0.f;
  if (!F.Q.ev_begin) return BNX_OK;       // untimed (bnx_render_noalloc)
  if (surfaces) {
    BNX_HIP(hipEventElapsedTime(&ms, F.Q.ev_begin, F.Q.ev_end));
    t.trace_ms += ms;
    t.trace_launches++;
  }
  if (volumes) {
    BNX_HIP(hipEventElapsedTime(&ms, F.Q.ev_vbegin, F.Q.ev_vend));
    t.march_ms += ms;
    t.march_launches++;
    t.march_rays += n;
  }
  return BNX_OK;
}

int read_stats(Device *dev, Wavefront &F, unsigned long long out[STAT_COUNT])
{
  unsigned long long *v = F.h_stats;
  BNX_HIP(hipMemcpyAsync(v, F.Q.stats.ptr, sizeof(F.h_stats), hipMemcpyDeviceToHost, dev->stream));
  BNX_HIP(hipStreamSynchronize(dev->stream));
  for (int k = 0; k < STAT_COUNT; k++) {
    out[k] = 0;
    for (int j = 0; j < STAT_SLOTS; j++) out[k] += v[(size_t)k * STAT_SLOTS + j];
  }
  return BNX_OK;
}

Wavefront &wavefront_for(Device *dev)
{
  if (!dev->queues) dev->queues = new Wavefront;
  return *(Wavefront *)dev->queues;
}

} // namespace

int trace_refill_for(const Device *dev, const World *w)
{
  return w->instances->blas_prims >= dev->trace_refill_prims ? dev->trace_refill : 64;
}

namespace {

// full: the march kernel the world needs (MARCH_FULL: it has isosurfaces or
// unstructured volumes; MARCH_CELLS: some unstructured field has other cells
// than tets)
int fill_slot(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, SlotDev &s, int &full)
{
  VolumeList vl{};
  int rc = world_volumes(w, vl);
  if (rc != BNX_OK) return rc;
  s = SlotDev{};
  s.scene = w->instances->scene;
  s.vols = vl.vols;
  s.num_vols = vl.count;
  if ((rc = world_isos(w, s.isos, s.num_isos)) != BNX_OK) return rc;
  full = s.num_isos > 0 ? MARCH_FULL : 0;
  for (const Volume *v : w->volumes) full |= march_kind(v);
  for (const IsoSurface *iso : w->isos) full |= march_kind(iso->field);
  BNX_SET(s.dir_lights, w->d_dir_lights.ptr);
  s.num_dir_lights = (int)w->dir_lights.size();
  s.env = env_dev(w);
  s.rr = rr;
  s.cam = cam;
  s.W = f->w;
  s.H = f->h;
  s.accum_base = f->accum_id;
  s.fb = frame_dev(f);
  return BNX_OK;
}

} // namespace

void wavefront_destroy(Device *dev)
{
  delete (Wavefront *)dev->queues;
  dev->queues = nullptr;
}

namespace {

// zero n 32-bit words: a kernel, where hipMemsetAsync would be a memset node
// in a caller's captured graph
__global__ void k_zero_u32(uint32_t *p, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

int zero_words(Device *dev, void *p, size_t bytes, bool kernel)
{
  if (!kernel) {
    BNX_HIP(hipMemsetAsync(p, 0, bytes, dev->stream));
    return BNX_OK;
  }
  const long long n = (long long)(bytes / 4);
  if (n > 0) k_zero_u32<<<(int)((n + 255) / 256), 256, 0, dev->stream>>>((uint32_t *)p, n);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

// the frame loop behind bnx_render and bnx_render_noalloc, over wavefront F;
// max_paths: paths per batch
int render_frame_impl(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, int spp,
                      bnx_render_stats *stats, Wavefront &F, long long max_paths)
{
  auto t0 = std::chrono::steady_clock::now();
  Device *dev = w->dev;
  const long long np = (long long)f->w * f->h;
  // samples per batch: about max_paths paths in flight
  const int S = (int)std::max(1LL, std::min({(long long)spp, max_paths / std::max(1LL, np),
                                             (long long)MAX_BATCH_SAMPLES}));
  BNX_CHECK(2 * np * S < (1LL << 31), BNX_ERR_UNSUPPORTED, "bnx_render: frame too large for one batch");
  int rc;
  // one batch of paths leaves at most twice as many rays (one shadow ray and
  // one continuation ray per path) in the output region
  if ((rc = F.ensure((size_t)(2 * np * S), 1, 1, 1)) != BNX_OK) return rc;
  if ((rc = zero_words(dev, F.Q.stats.ptr, sizeof(unsigned long long) * STAT_COUNT * STAT_SLOTS, F.noalloc)) !=
      BNX_OK)
    return rc;
  SlotDev *hs = F.h_slots.as<SlotDev>();
  int full = 0;
  if ((rc = fill_slot(w, cam, rr, f, *hs, full)) != BNX_OK) return rc;
  const bool surfaces = hs->scene.num_insts > 0;
  const bool volumes = hs->num_vols > 0 || hs->num_isos > 0;
  const bool traced = surfaces || !volumes;
  RaygenChunk *hc = F.h_chunks.as<RaygenChunk>();
  int *hb = F.h_out_base.as<int>();
  int *hcnt = F.h_out_cnt.as<int>();
  hb[0] = 0;                           // the single request's region starts the queue
  BNX_HIP(hipMemcpyAsync(F.out_base.ptr, hb, sizeof(int), hipMemcpyHostToDevice, dev->stream));
  int64_t rays = 0, gens = 0, captured = 0, uncaptured = 0;
  GenTiming tm;
  for (int s0 = 0; s0 < spp; s0 += S) {
    const int nS = std::min(S, spp - s0);
    hs->accum_base = f->accum_id;
    hc[0] = RaygenChunk{0, 0, 0, 0};
    BNX_HIP(hipMemcpyAsync(F.slots.ptr, hs, sizeof(SlotDev), hipMemcpyHostToDevice, dev->stream));
    BNX_HIP(hipMemcpyAsync(F.chunks.ptr, hc, sizeof(RaygenChunk), hipMemcpyHostToDevice, dev->stream));
    int cur = 0;
    int count = (int)(np * nS);
    if ((rc = launch_raygen(dev, F.chunks.as<RaygenChunk>(), 1, F.slots.as<SlotDev>(), 0, count, F.Q.q(cur))) !=
        BNX_OK)
      return rc;
    bool primary = true;
    // batches of up to graph_paths paths: once the generations a batch of
    // this shape takes are known (from a first batch through the host loop
    // below), run that many as one captured bounce loop with one read-back
    const bool dd = !F.noalloc && dev->graph_paths > 0 && np * nS <= dev->graph_paths;
    const BatchShape shape{np * nS,          2 * np * nS,           surfaces,
                           volumes,          full,                  trace_refill_for(dev, w),
                           dev->trace_stack_lds, dev->trace_walk, dev->trace_blocks_per_cu, dev->shade_block,
                           ext_kernels(dev, has_cut_plane(rr)), dev->xf_worlds > 0};
    int dd_gens = 0;
    if (dd) {
      auto it = F.gens_seen.find(shape);
      if (it != F.gens_seen.end()) {
        const int G = it->second;
        if ((rc = F.ensure_dd()) != BNX_OK) return rc;
        hipGraphExec_t exec = nullptr;
        if ((rc = batch_graph(dev, F, shape, G, cur, exec)) != BNX_OK) return rc;
        if (exec) {
          BNX_HIP(hipGraphLaunch(exec, dev->stream));
          captured++;
        } else if ((rc = enqueue_dd(dev, F, cur, G, shape)) != BNX_OK) {
          return rc;
        } else {
          uncaptured++;
        }
        long long *hg = F.h_gen_counts.as<long long>();
        unsigned long long *hts = (unsigned long long *)(hg + MAX_DD_GENS);
        int *hc2 = (int *)(hg + 5 * MAX_DD_GENS);
        BNX_HIP(hipMemcpyAsync(hg, F.gen_counts.ptr, sizeof(long long) * G, hipMemcpyDeviceToHost, dev->stream));
        BNX_HIP(hipMemcpyAsync(hts, F.stamps.ptr, sizeof(unsigned long long) * 4 * G, hipMemcpyDeviceToHost,
                               dev->stream));
        BNX_HIP(hipMemcpyAsync(hc2, F.out_cnt.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, dev->stream));
        BNX_HIP(hipStreamSynchronize(dev->stream));
        const double tick_ms = 1.0 / (double)dev->wall_clock_khz;
        for (int g = 0; g < G; g++) {
          if (traced) {
            tm.trace_ms += (double)(hts[4 * g + 1] - hts[4 * g]) * tick_ms;
            tm.trace_launches++;
          }
          if (volumes) {
            tm.march_ms += (double)(hts[4 * g + 3] - hts[4 * g + 2]) * tick_ms;
            tm.march_launches++;
            tm.march_rays += hg[g];
          }
          rays += hg[g];
          gens += hg[g] > 0;
        }
        count = hc2[0];
        cur = (cur + G) & 1;
        primary = false;
        dd_gens = G;
      }
    }
    // the host loop: one read-back per generation (and the generations a
    // captured loop left over)
    int host_gens = 0;
    // secondary generations of surface worlds may be reordered first (into
    // the other queue, whose rays are spent)
    const Bvh &tlas = w->instances->bvh;
    const bool sorting = !F.noalloc && dev->sort_rays != 0 && traced && tlas.num_prims > 0;
    for (; count > 0; primary = false) {
      if (sorting && !primary && count >= SORT_MIN_RAYS) {
        if ((rc = sort_rays(dev, F.sorter, F.Q.q(cur), F.Q.q(1 - cur), count, tlas.root_lo, tlas.root_hi)) != BNX_OK)
          return rc;
        cur = 1 - cur;
      }
      SegList sl{F.h_segs.as<Segment>()};
      sl.add(0, 0, count, primary ? 64 : trace_refill_for(dev, w));
      if ((rc = zero_words(dev, F.out_cnt.ptr, 2 * sizeof(int), F.noalloc)) != BNX_OK) return rc;
      if ((rc = enqueue_generation(dev, F, cur, sl, surfaces, volumes, full)) != BNX_OK) return rc;
      BNX_HIP(hipMemcpyAsync(hcnt, F.out_cnt.ptr, 2 * sizeof(int), hipMemcpyDeviceToHost, dev->stream));
      BNX_HIP(hipStreamSynchronize(dev->stream));
      if ((rc = collect_timing(F, count, traced, volumes, tm)) != BNX_OK) return rc;
      rays += count;
      gens++;
      host_gens++;
      count = hcnt[0];
      cur = 1 - cur;
    }
    if (dd && host_gens > 0) {
      if (F.gens_seen.size() >= 256) F.gens_seen.clear();
      F.gens_seen[shape] = std::min(MAX_DD_GENS, dd_gens + host_gens);
    }
    f->accum_id += nS;
  }
  unsigned long long host_st[STAT_COUNT];
  if ((rc = read_stats(dev, F, host_st)) != BNX_OK) return rc;
  BNX_HIP(hipGetLastError());
  if (stats) {
    stats->rays_traced = rays;
    stats->shadow_rays = (int64_t)host_st[STAT_SHADOW];
    stats->generations = gens;
    stats->volume_samples = (int64_t)host_st[STAT_SAMPLES];
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->node_visits = (int64_t)host_st[STAT_NODES];
    stats->tri_tests = (int64_t)host_st[STAT_TRIS];
    stats->sphere_tests = (int64_t)host_st[STAT_SPHERES];
    stats->trace_ms = tm.trace_ms;
    stats->trace_launches = tm.trace_launches;
    stats->march_ms = tm.march_ms;
    stats->march_launches = tm.march_launches;
    stats->mc_cells = (int64_t)host_st[STAT_CELLS];
    stats->march_rays = tm.march_rays;
    stats->umesh_nodes = (int64_t)host_st[STAT_UNODES];
    stats->umesh_tets = (int64_t)host_st[STAT_UTETS];
    stats->captured_batches = captured;
    stats->uncaptured_batches = uncaptured;
    stats->capture_error = (int32_t)F.capture_err;
    stats->reserved = 0;
  }
  return BNX_OK;
}

} // namespace

int render_frame(World *w, const bnx_camera &cam, const bnx_renderer &rr, Frame *f, int spp, bnx_render_stats *stats)
{
  Device *dev = w->dev;
  return render_frame_impl(w, cam, rr, f, spp, stats, wavefront_for(dev), dev->render_paths);
}

// ------------------------------------------------------ bnx_render_noalloc
namespace {

// Paths per batch of bnx_render_noalloc: the queues hold twice as many rays
// (one shadow and one continuation ray per path), two queues of 92 B per
// ray: 24 MB of the static array below.
constexpr long long NOALLOC_PATHS = 1LL << 16;
constexpr size_t NOALLOC_RAYS = 2 * (size_t)NOALLOC_PATHS;
constexpr size_t NOALLOC_BLOCKS = NOALLOC_RAYS / 256 + 2;   // one segment: one partial block at most
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_QUEUE_BYTES =
    al256(16 * NOALLOC_RAYS) * 5 + al256(8 * NOALLOC_RAYS) + al256(4 * NOALLOC_RAYS);   // one queue
constexpr size_t NOALLOC_BYTES = 2 * NOALLOC_QUEUE_BYTES + al256(8 * STAT_COUNT * STAT_SLOTS) +
                                 al256(sizeof(SlotDev)) + al256(sizeof(RaygenChunk)) + al256(sizeof(Segment)) +
                                 4 * 256 + 3 * al256(4 * NOALLOC_BLOCKS);

// The scratch memory of bnx_render_noalloc: one copy per GPU the library
// runs on, held until the process ends.  Host code never passes this name to
// a kernel or copy: it uses the address hipGetSymbolAddress gives for the
// current device (Device::noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// lay the noalloc wavefront's buffers out in the scratch array at dev_base
// and in its host staging
struct NoallocWavefront {
  Wavefront F;
  unsigned char host_stage[4096];
};
static_assert(sizeof(SlotDev) + sizeof(RaygenChunk) + sizeof(Segment) + 64 <= 4096, "host staging");

void noalloc_layout(NoallocWavefront &N, unsigned char *base)
{
  Wavefront &F = N.F;
  size_t o = 0;
  auto take = [&](DevMem &m, size_t b) {
    m.borrow(base + o, b);
    o += al256(b);
  };
  for (int k = 0; k < 2; k++) {
    take(F.Q.org[k], 16 * NOALLOC_RAYS);
    take(F.Q.dir[k], 16 * NOALLOC_RAYS);
    take(F.Q.rng[k], 8 * NOALLOC_RAYS);
    take(F.Q.ps[k], 16 * NOALLOC_RAYS);
    take(F.Q.hit[k], 16 * NOALLOC_RAYS);
    take(F.Q.hit_v[k], 4 * NOALLOC_RAYS);
    take(F.Q.vcol[k], 16 * NOALLOC_RAYS);
  }
  take(F.Q.stats, 8 * STAT_COUNT * STAT_SLOTS);
  take(F.slots, sizeof(SlotDev));
  take(F.chunks, sizeof(RaygenChunk));
  take(F.segs, sizeof(Segment));
  take(F.pool, 4);
  take(F.out_base, 4);
  take(F.out_cnt, 8);
  take(F.blk256, 4 * NOALLOC_BLOCKS);
  take(F.blk1024, 4 * NOALLOC_BLOCKS);
  take(F.blkT, 4 * NOALLOC_BLOCKS);
  F.Q.cap = NOALLOC_RAYS;
  F.slot_cap = F.chunk_cap = F.seg_cap = 1;
  F.blk_cap = NOALLOC_BLOCKS;
  size_t h = 0;
  auto host = [&](Pinned &p, size_t b) {
    p.borrow(N.host_stage + h, b);
    h += (b + 15) / 16 * 16;
  };
  host(F.h_slots, sizeof(SlotDev));
  host(F.h_chunks, sizeof(RaygenChunk));
  host(F.h_segs, sizeof(Segment));
  host(F.h_out_base, 4);
  host(F.h_out_cnt, 8);
  F.noalloc = true;
  F.graphs_off = true;
}

} // namespace

int noalloc_init(Device *dev)
{
  void *p = nullptr;
  BNX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  dev->noalloc_scratch = p;
  NoallocWavefront *N = new NoallocWavefront;
  noalloc_layout(*N, (unsigned char *)p);
  dev->noalloc_queues = N;
  return BNX_OK;
}

void noalloc_destroy(Device *dev)
{
  delete (NoallocWavefront *)dev->noalloc_queues;
  dev->noalloc_queues = nullptr;
  dev->noalloc_scratch = nullptr;
}

// ------------------------------------------------------------------ server
namespace {

constexpr size_t FRAME_BYTES_PER_PIXEL = 80;   // 64 B of channels + 16 B read staging

// ---- device-driven server steps: several generations per host read-back.
// A request slot's state as the device keeps it through such a step; the
// host uploads it for every active slot before the step and reads it back
// after it.
struct SrvSlot {
  long long next;              // paths started
  long long total;             // paths of the request
  long long rays;              // rays of the slot that entered a generation in this step
  int region;                  // first queue index of its rays in the current input queue
  int live_rays, live_paths;   // rays there; paths they continue
  int refill;                  // trace-pool refill of its secondary rays
  int added;                   // primary rays added in the current generation
  int pad;
};
static_assert(sizeof(SrvSlot) == 48, "SrvSlot");

// one step's tables: the active slots in the host's order (their old regions
// come first in every generation), the slots with primaries left in top-up
// priority order, and the scalars the generations hand to each other
struct SrvStep {
  SrvSlot *st;
  const int *act;
  const int *ord;
  int A, O;
  Segment *segs;               // A + O: old regions, then new primaries
  RaygenChunk *chunks;         // O
  int *out_base, *out_cnt;
  // [0] end of the regions in the current input queue (new primaries go
  // there), [1] [2] the raygen range (queue begin, count) of this generation
  long long *scal;
  long long *gen_counts;       // rays per generation
  long long max_paths;
  int target;                  // trace blocks to aim for (pooled layout)
};

constexpr int SRV_TB = 1024;

// exclusive prefix sum over the SRV_TB threads of the block; total: the sum
__device__ long long srv_scan(long long v, long long &total)
{
  __shared__ long long wsum[SRV_TB / 64];
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  long long x = v;
  for (int d = 1; d < 64; d <<= 1) {
    const long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  long long off = 0, tot = 0;
  for (int k = 0; k < SRV_TB / 64; k++) {
    const long long s = wsum[k];
    off += k < w ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

__device__ long long srv_sum(long long v)
{
  long long t;
  (void)srv_scan(v, t);
  return t;
}

// The host step's bookkeeping for one generation, on the device (one block
// of SRV_TB threads):
//  - fold the previous generation's output (g > 0): each slot's rays are now
//    the region its shade wrote, with the counts it wrote;
//  - top up: walk the slots with primaries left in priority order, each
//    taking min(what is left, room) paths while the paths in flight stay
//    within max_paths; their raygen chunks follow the old regions;
//  - segments: old regions of every active slot, then the new primaries,
//    with their block offsets for both launch widths and the trace pool
//    layout (about `target` blocks, as SegList::layout_trace);
//  - output regions in the other queue, 2 rays per path in flight, in the
//    order of the active slots, with their counters cleared.
// fold_only: the fold alone (after a step's last generation).
__global__ __launch_bounds__(SRV_TB) void k_srv_setup(SrvStep p, int g, int fold_only)
{
  const int tid = (int)threadIdx.x;
  if (g > 0 || fold_only) {
    for (int i = tid; i < p.A; i += SRV_TB) {
      const int s = p.act[i];
      p.st[s].region = p.out_base[s];
      p.st[s].live_rays = p.out_cnt[2 * s];
      p.st[s].live_paths = p.out_cnt[2 * s + 1];
    }
  }
  if (fold_only) return;
  for (int i = tid; i < p.A; i += SRV_TB) p.st[p.act[i]].added = 0;
  __syncthreads();
  const long long in_end = p.scal[0];
  long long lp = 0;
  for (int i = tid; i < p.A; i += SRV_TB) lp += p.st[p.act[i]].live_paths;
  const long long room = max(0LL, p.max_paths - srv_sum(lp));
  // top-up: the takes are the remaining paths clamped to what room is left
  // after the slots before, so their prefix sum is min(prefix of rem, room)
  long long carry = 0;
  for (int base = 0; base < p.O; base += SRV_TB) {
    const int j = base + tid;
    int s = -1;
    long long rem = 0;
    if (j < p.O) {
      s = p.ord[j];
      rem = p.st[s].total - p.st[s].next;
    }
    long long tot;
    const long long pre = carry + srv_scan(rem, tot);
    if (j < p.O) {
      const long long take = min(rem, max(0LL, room - pre));
      p.chunks[j] = RaygenChunk{in_end + min(pre, room), p.st[s].next, s, 0};
      p.st[s].next += take;
      p.st[s].added = (int)take;
    }
    carry += tot;
  }
  const long long added = min(carry, room);
  __syncthreads();
  // segments
  const int nsegs = p.A + p.O;
  auto seg = [&](int k, int &slot, int &begin, int &refill) -> int {
    if (k < p.A) {
      slot = p.act[k];
      begin = p.st[slot].region;
      refill = p.st[slot].refill;
      return p.st[slot].live_rays;
    }
    slot = p.ord[k - p.A];
    begin = (int)p.chunks[k - p.A].begin;
    refill = 64;                 // primary rays: whole-wave refills
    return p.st[slot].added;
  };
  long long n = 0;
  for (int k = tid; k < nsegs; k += SRV_TB) {
    int slot, begin, refill;
    n += seg(k, slot, begin, refill);
  }
  n = srv_sum(n);
  const long long per = 256LL * max(1LL, (n + 256LL * p.target - 1) / (256LL * p.target));
  long long c256 = 0, c1024 = 0, cT = 0;
  for (int base = 0; base < nsegs; base += SRV_TB) {
    const int k = base + tid;
    int slot = 0, begin = 0, refill = 64, count = 0;
    if (k < nsegs) count = seg(k, slot, begin, refill);
    long long t256, t1024, tT;
    const long long b256 = c256 + srv_scan((count + 255) / 256, t256);
    const long long b1024 = c1024 + srv_scan((count + 1023) / 1024, t1024);
    const long long bT = cT + srv_scan((count + per - 1) / per, tT);
    if (k < nsegs) p.segs[k] = Segment{slot, begin, count, (int)b256, (int)b1024, (int)bT, refill, 0};
    c256 += t256;
    c1024 += t1024;
    cT += tT;
  }
  // output regions
  carry = 0;
  for (int base = 0; base < p.A; base += SRV_TB) {
    const int i = base + tid;
    long long size = 0;
    int s = -1;
    if (i < p.A) {
      s = p.act[i];
      size = 2LL * (p.st[s].live_paths + p.st[s].added);
    }
    long long tot;
    const long long pre = carry + srv_scan(size, tot);
    if (i < p.A) {
      p.out_base[s] = (int)pre;
      p.out_cnt[2 * s] = 0;
      p.out_cnt[2 * s + 1] = 0;
      p.st[s].rays += p.st[s].live_rays + p.st[s].added;
    }
    carry += tot;
  }
  if (tid == 0) {
    p.gen_counts[g] = n;
    p.scal[1] = in_end;
    p.scal[2] = added;
    p.scal[0] = carry;           // the next generation's input regions end here
  }
}

struct Request {
  int64_t ticket = 0;
  World *world = nullptr;
  bnx_camera cam{};
  bnx_renderer rr{};
  int W = 0, H = 0, spp = 0;
  int state = BNX_REQ_QUEUED;
  int slot = -1;
  Frame *frame = nullptr;
  long long next_path = 0, total_paths = 0;
  int region = 0;                  // first queue index of its rays in the current queue
  int live_rays = 0, live_paths = 0;
  int added = 0;                   // primary rays added in this step
  double submit_ms = -1.0, admit_ms = -1.0, done_ms = -1.0;
  double cost_ms = 0.0;            // share of step time by its rays in each step
  double deadline_ms = -1.0;       // set at admission when the top-up is deadline-ordered
  bool surfaces = false, volumes = false;
  int full = 0;
};

// measured cost of a world's paths: step time attributed to completed
// requests by their share of each generation's rays, over their paths
struct PathCost {
  double ms = 0.0, paths = 0.0;
};

} // namespace

struct Server {
  Device *dev = nullptr;
  long long max_paths = 0;
  size_t budget = 0;
  int max_active = 0;
  float slack = 0.f;               // > 0: earliest-deadline-first top-up; 0: shortest remaining first
  std::unordered_map<const World *, PathCost> world_cost;
  PathCost all_cost;
  Wavefront F;
  std::unordered_map<int64_t, Request *> reqs;
  std::deque<Request *> queued;
  std::vector<Request *> active;
  std::vector<int> free_slots;
  std::multimap<std::pair<int, int>, Frame *> pool;
  size_t frame_bytes = 0;
  int cur = 0;
  int in_end = 0;                  // end of the output regions laid out in the current queue
  // device-driven steps: slot states, active and top-up order, scalars
  DevMem d_srv, d_act, d_ord, d_scal;
  Pinned h_srv, h_act, h_ord, h_scal;
  int64_t next_ticket = 1;
  int64_t submitted = 0, admitted = 0, done = 0, released = 0, failed = 0, rays = 0, gens = 0, peak = 0;
  double busy_ms = 0.0;
  GenTiming tm;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now_ms() const
  {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  ~Server()
  {
    (void)hipStreamSynchronize(dev->stream);
    for (auto &kv : reqs) {
      delete kv.second->frame;
      delete kv.second;
    }
    for (auto &kv : pool) delete kv.second;
  }

  static size_t bytes_of(int w, int h) { return FRAME_BYTES_PER_PIXEL * (size_t)w * h; }
  // a cleared frame of w x h from the pool or newly allocated; null if the
  // budget does not allow it right now
  int take_frame(int w, int h, Frame **out)
  {
    *out = nullptr;
    auto it = pool.find({w, h});
    if (it != pool.end()) {
      *out = it->second;
      pool.erase(it);
    } else {
      const size_t need = bytes_of(w, h);
      while (frame_bytes + need > budget && !pool.empty()) {
        auto v = pool.begin();
        frame_bytes -= bytes_of(v->second->w, v->second->h);
        delete v->second;
        pool.erase(v);
      }
      if (frame_bytes + need > budget) return BNX_OK;
      Frame *f = new Frame;
      f->dev = dev;
      f->w = w;
      f->h = h;
      const size_t np = (size_t)w * h;
      hipError_t e;
      if ((e = f->accum.ensure(32 * np)) != hipSuccess || (e = f->depth.ensure(4 * np)) != hipSuccess ||
          (e = f->prim.ensure(4 * np)) != hipSuccess || (e = f->inst.ensure(4 * np)) != hipSuccess ||
          (e = f->obj.ensure(4 * np)) != hipSuccess || (e = f->normal.ensure(16 * np)) != hipSuccess ||
          (e = f->out.ensure(16 * np)) != hipSuccess) {
        delete f;
        set_error(std::string("bnx_server_step: frame allocation: ") + hipGetErrorString(e));
        return BNX_ERR_OOM;
      }
      frame_bytes += need;
      *out = f;
    }
    return frame_clear(*out);
  }

  // a request's time alone estimated from its world's measured cost per
  // path (all worlds' when none of its world completed yet; 0 before any)
  double estimate_ms(const Request *r) const
  {
    auto it = world_cost.find(r->world);
    const PathCost &c = (it != world_cost.end() && it->second.paths > 0.0) ? it->second : all_cost;
    return c.paths > 0.0 ? (double)r->total_paths * (c.ms / c.paths) : 0.0;
  }

  int admit()
  {
    while (!queued.empty() && !free_slots.empty()) {
      Request *r = queued.front();
      Frame *f = nullptr;
      int rc = take_frame(r->W, r->H, &f);
      if (rc != BNX_OK) return rc;
      if (!f) break;                       // admission is in arrival order
      queued.pop_front();
      r->frame = f;
      r->slot = free_slots.back();
      free_slots.pop_back();
      SlotDev *hs = F.h_slots.as<SlotDev>() + r->slot;
      if ((rc = fill_slot(r->world, r->cam, r->rr, f, *hs, r->full)) != BNX_OK) {
        // the request fails: its slot and its frame go back, it leaves the
        // queue, and its ticket reads BNX_REQ_FAILED until released
        free_slots.push_back(r->slot);
        r->slot = -1;
        pool.insert({{f->w, f->h}, f});
        r->frame = nullptr;
        r->state = BNX_REQ_FAILED;
        r->done_ms = now_ms();
        failed++;
        return rc;
      }
      hs->accum_base = 0;
      BNX_HIP(hipMemcpyAsync(F.slots.as<SlotDev>() + r->slot, hs, sizeof(SlotDev), hipMemcpyHostToDevice,
                             dev->stream));
      r->surfaces = hs->scene.num_insts > 0;
      r->volumes = hs->num_vols > 0 || hs->num_isos > 0;
      r->state = BNX_REQ_ACTIVE;
      r->admit_ms = now_ms();
      r->deadline_ms = r->submit_ms + slack * estimate_ms(r);
      active.push_back(r);
      admitted++;
    }
    return BNX_OK;
  }

  // the active requests with primaries left, in top-up priority order
  std::vector<Request *> topup_order() const
  {
    std::vector<Request *> order;
    for (Request *r : active)
      if (r->next_path < r->total_paths) order.push_back(r);
    if (slack > 0.f)
      std::stable_sort(order.begin(), order.end(),
                       [](const Request *a, const Request *b) { return a->deadline_ms < b->deadline_ms; });
    else
      std::stable_sort(order.begin(), order.end(), [](const Request *a, const Request *b) {
        return a->total_paths - a->next_path < b->total_paths - b->next_path;
      });
    return order;
  }

  // after a step: a request whose primaries are all issued and that has no
  // rays left is done; its slot is freed
  void settle(double tnow, int64_t *done_tickets, int max_done, int *num_done)
  {
    std::vector<Request *> still;
    for (Request *r : active) {
      if (r->next_path == r->total_paths && r->live_rays == 0) {
        r->state = BNX_REQ_DONE;
        r->done_ms = tnow;
        PathCost &wc = world_cost[r->world];
        wc.ms += r->cost_ms;
        wc.paths += (double)r->total_paths;
        all_cost.ms += r->cost_ms;
        all_cost.paths += (double)r->total_paths;
        r->frame->accum_id = r->spp;
        free_slots.push_back(r->slot);
        r->slot = -1;
        done++;
        if (*num_done < max_done && done_tickets) done_tickets[*num_done] = r->ticket;
        (*num_done)++;
      } else {
        still.push_back(r);
      }
    }
    active.swap(still);
  }

  // One generation for every request in flight.  The queue holds, per
  // active request, the rays its last shade wrote (region, live_rays); new
  // primary rays go after the regions, shortest remaining work first, while
  // the paths in flight stay within max_paths.  The next output regions are
  // laid out in the other queue, 2 rays per path.
  int step(int64_t *done_tickets, int max_done, int *num_done)
  {
    const double ts = now_ms();
    *num_done = 0;
    int rc;
    if ((rc = admit()) != BNX_OK) return rc;
    if (active.empty()) return BNX_OK;
    long long in_flight = 0;
    for (Request *r : active) {
      in_flight += r->live_paths;
      r->added = 0;
    }
    const std::vector<Request *> order = topup_order();
    long long room = max_paths - in_flight;
    int nch = 0;
    long long added = 0;
    RaygenChunk *hc = F.h_chunks.as<RaygenChunk>();
    // this generation's segments: old regions, then the new primaries
    SegList sl{F.h_segs.as<Segment>()};
    for (Request *r : active) sl.add(r->slot, r->region, r->live_rays, trace_refill_for(dev, r->world));
    for (Request *r : order) {
      if (room <= 0) break;
      const long long take = std::min(room, r->total_paths - r->next_path);
      hc[nch++] = RaygenChunk{(long long)in_end + added, r->next_path, r->slot, 0};
      sl.add(r->slot, (int)(in_end + added), (int)take, 64);
      r->added = (int)take;
      r->next_path += take;
      added += take;
      room -= take;
    }
    long long n = 0;
    for (int k = 0; k < sl.n; k++) n += sl.v[k].count;
    if (n == 0) {
      busy_ms += now_ms() - ts;
      return BNX_OK;
    }
    // output regions in the other queue
    int *hb = F.h_out_base.as<int>();
    long long out_end = 0;
    bool surfaces = false, volumes = false;
    int full = 0;
    int max_slot = 0;
    for (Request *r : active) {
      hb[r->slot] = (int)out_end;
      out_end += 2LL * (r->live_paths + r->added);
      surfaces |= r->surfaces;
      volumes |= r->volumes;
      full |= r->full;
      max_slot = std::max(max_slot, r->slot + 1);
    }
    BNX_HIP(hipMemcpyAsync(F.out_base.ptr, hb, sizeof(int) * max_slot, hipMemcpyHostToDevice, dev->stream));
    BNX_HIP(hipMemsetAsync(F.out_cnt.ptr, 0, sizeof(int) * 2 * max_slot, dev->stream));
    if (nch > 0) {
      BNX_HIP(hipMemcpyAsync(F.chunks.ptr, hc, sizeof(RaygenChunk) * nch, hipMemcpyHostToDevice, dev->stream));
      if ((rc = launch_raygen(dev, F.chunks.as<RaygenChunk>(), nch, F.slots.as<SlotDev>(), in_end, added,
                              F.Q.q(cur))) != BNX_OK)
        return rc;
    }
    if ((rc = enqueue_generation(dev, F, cur, sl, surfaces, volumes, full)) != BNX_OK) return rc;
    int *hcnt = F.h_out_cnt.as<int>();
    BNX_HIP(hipMemcpyAsync(hcnt, F.out_cnt.ptr, sizeof(int) * 2 * max_slot, hipMemcpyDeviceToHost, dev->stream));
    BNX_HIP(hipStreamSynchronize(dev->stream));
    BNX_HIP(hipGetLastError());
    if ((rc = collect_timing(F, (int)n, surfaces || !volumes, volumes, tm)) != BNX_OK) return rc;
    rays += n;
    gens++;
    peak = std::max<int64_t>(peak, n);
    cur = 1 - cur;
    in_end = (int)out_end;
    const double tnow = now_ms();
    const double per_ray = (tnow - ts) / (double)n;
    for (Request *r : active) {
      r->cost_ms += per_ray * (double)(r->live_rays + r->added);
      r->region = hb[r->slot];
      r->live_rays = hcnt[2 * r->slot];
      r->live_paths = hcnt[2 * r->slot + 1];
    }
    settle(tnow, done_tickets, max_done, num_done);
    busy_ms += now_ms() - ts;
    return BNX_OK;
  }

  int ensure_dd()
  {
    const size_t act = (size_t)max_active;
    BNX_HIP(d_srv.ensure(sizeof(SrvSlot) * act));
    BNX_HIP(h_srv.ensure(sizeof(SrvSlot) * act));
    BNX_HIP(d_act.ensure(sizeof(int) * act));
    BNX_HIP(h_act.ensure(sizeof(int) * act));
    BNX_HIP(d_ord.ensure(sizeof(int) * act));
    BNX_HIP(h_ord.ensure(sizeof(int) * act));
    BNX_HIP(d_scal.ensure(sizeof(long long) * 4));
    BNX_HIP(h_scal.ensure(sizeof(long long) * 4));
    return F.ensure_dd();
  }

  // G generations for every request in flight with the step's bookkeeping
  // on the device (k_srv_setup before each generation) and one read-back at
  // the end: admission happens before the first generation only, the top-up
  // runs before every one.  Launches are sized for the largest generation
  // the step can hold (blocks past a generation's rays find none), from
  // bounds the host knows: the first generation's old rays exactly, then at
  // most two rays per path in flight, P = min(max_paths, paths in flight +
  // primaries left) paths, plus the primaries a top-up can add.
  int step_dd(int G, int64_t *done_tickets, int max_done, int *num_done)
  {
    const double ts = now_ms();
    *num_done = 0;
    int rc;
    if ((rc = admit()) != BNX_OK) return rc;
    if (active.empty()) return BNX_OK;
    if ((rc = ensure_dd()) != BNX_OK) return rc;
    const std::vector<Request *> order = topup_order();
    SrvSlot *hs = h_srv.as<SrvSlot>();
    int *ha = h_act.as<int>(), *ho = h_ord.as<int>();
    long long *hsc = h_scal.as<long long>();
    const int A = (int)active.size(), O = (int)order.size();
    long long rays0 = 0, paths = 0, rem = 0;
    bool surfaces = false, volumes = false, pooled = false, cut = false;
    int full = 0;
    int max_slot = 0;
    for (int i = 0; i < A; i++) {
      Request *r = active[i];
      const int refill = trace_refill_for(dev, r->world);
      ha[i] = r->slot;
      hs[r->slot] = SrvSlot{r->next_path, r->total_paths, 0, r->region, r->live_rays, r->live_paths, refill, 0, 0};
      rays0 += r->live_rays;
      paths += r->live_paths;
      rem += r->total_paths - r->next_path;
      surfaces |= r->surfaces;
      volumes |= r->volumes;
      full |= r->full;
      pooled |= refill < 64;
      cut |= has_cut_plane(r->rr);
      max_slot = std::max(max_slot, r->slot + 1);
    }
    for (int j = 0; j < O; j++) ho[j] = order[j]->slot;
    hsc[0] = in_end;
    hipStream_t st = dev->stream;
    BNX_HIP(hipMemcpyAsync(d_srv.ptr, hs, sizeof(SrvSlot) * max_slot, hipMemcpyHostToDevice, st));
    BNX_HIP(hipMemcpyAsync(d_act.ptr, ha, sizeof(int) * A, hipMemcpyHostToDevice, st));
    if (O > 0) BNX_HIP(hipMemcpyAsync(d_ord.ptr, ho, sizeof(int) * O, hipMemcpyHostToDevice, st));
    BNX_HIP(hipMemcpyAsync(d_scal.ptr, hsc, sizeof(long long), hipMemcpyHostToDevice, st));
    const long long added_max = std::min(max_paths, rem);
    const long long P = std::min(max_paths, paths + rem);
    const int nsegs = A + O;
    const int target = dev->trace_blocks_per_cu * dev->num_cus;
    const SrvStep p{d_srv.as<SrvSlot>(), d_act.as<int>(), d_ord.as<int>(), A, O, F.segs.as<Segment>(),
                    F.chunks.as<RaygenChunk>(), F.out_base.as<int>(), F.out_cnt.as<int>(),
                    d_scal.as<long long>(), F.gen_counts.as<long long>(), max_paths, target};
    const SegView sv{F.segs.as<Segment>(), F.blk256.as<int>(), F.blk1024.as<int>(), F.blkT.as<int>(),
                     F.pool.as<int>()};
    const SlotOut so{F.out_base.as<int>(), F.out_cnt.as<int>()};
    const SlotDev *slots = F.slots.as<SlotDev>();
    unsigned long long *stat = F.Q.stats.as<unsigned long long>();
    const bool trace = surfaces || !volumes;
    for (int g = 0; g < G; g++) {
      const long long N = g == 0 ? rays0 + added_max : 2 * P + added_max;
      const int blocks256 = (int)((N + 255) / 256) + nsegs, blocks1024 = (int)((N + 1023) / 1024) + nsegs;
      const int blocksT = pooled ? std::min(target + nsegs, blocks256) : 0;
      k_srv_setup<<<1, SRV_TB, 0, st>>>(p, g, 0);
      BNX_HIP(hipGetLastError());
      RayQueue in = F.Q.q(cur), out = F.Q.q(1 - cur);
      if (added_max > 0 && O > 0 &&
          (rc = launch_raygen_dev(dev, p.chunks, O, slots, p.scal + 1, added_max, in)) != BNX_OK)
        return rc;
      if ((rc = launch_segment_blocks(dev, sv, nsegs, blocks256, blocks1024, blocksT)) != BNX_OK) return rc;
      TraceQueue tq{in.org, in.dir, in.hit, in.hit_v, in.rng, in.ps};
      unsigned long long *tsp = F.stamps.as<unsigned long long>() + 4 * g;
      if (trace) {
        if ((rc = launch_stamp(dev, tsp)) != BNX_OK) return rc;
        if ((rc = launch_trace_queue(dev, slots, sv, blocks256, blocksT, pooled, cut, tq, stat)) != BNX_OK)
          return rc;
        if ((rc = launch_stamp(dev, tsp + 1)) != BNX_OK) return rc;
      }
      if (volumes) {
        if ((rc = launch_stamp(dev, tsp + 2)) != BNX_OK) return rc;
        if ((rc = launch_volume_march(dev, slots, sv, blocks256, in, trace, full, stat)) != BNX_OK) return rc;
        if ((rc = launch_stamp(dev, tsp + 3)) != BNX_OK) return rc;
      }
      if ((rc = launch_shade(dev, slots, sv, blocks256, blocks1024, in, out, so, stat)) != BNX_OK) return rc;
      cur = 1 - cur;
    }
    k_srv_setup<<<1, SRV_TB, 0, st>>>(p, G, 1);
    BNX_HIP(hipGetLastError());
    long long *hg = F.h_gen_counts.as<long long>();
    unsigned long long *hts = (unsigned long long *)(hg + MAX_DD_GENS);
    BNX_HIP(hipMemcpyAsync(hs, d_srv.ptr, sizeof(SrvSlot) * max_slot, hipMemcpyDeviceToHost, st));
    BNX_HIP(hipMemcpyAsync(hsc, d_scal.ptr, sizeof(long long), hipMemcpyDeviceToHost, st));
    BNX_HIP(hipMemcpyAsync(hg, F.gen_counts.ptr, sizeof(long long) * G, hipMemcpyDeviceToHost, st));
    BNX_HIP(hipMemcpyAsync(hts, F.stamps.ptr, sizeof(unsigned long long) * 4 * G, hipMemcpyDeviceToHost, st));
    BNX_HIP(hipStreamSynchronize(st));
    BNX_HIP(hipGetLastError());
    const double tick_ms = 1.0 / (double)dev->wall_clock_khz;
    long long n = 0;
    for (int g = 0; g < G; g++) {
      if (trace) {
        tm.trace_ms += (double)(hts[4 * g + 1] - hts[4 * g]) * tick_ms;
        tm.trace_launches++;
      }
      if (volumes) {
        tm.march_ms += (double)(hts[4 * g + 3] - hts[4 * g + 2]) * tick_ms;
        tm.march_launches++;
        tm.march_rays += hg[g];
      }
      n += hg[g];
      gens += hg[g] > 0;
      peak = std::max<int64_t>(peak, hg[g]);
    }
    rays += n;
    in_end = (int)hsc[0];
    const double tnow = now_ms();
    const double per_ray = n > 0 ? (tnow - ts) / (double)n : 0.0;
    for (Request *r : active) {
      const SrvSlot &q = hs[r->slot];
      r->cost_ms += per_ray * (double)q.rays;
      r->next_path = q.next;
      r->region = q.region;
      r->live_rays = q.live_rays;
      r->live_paths = q.live_paths;
    }
    settle(tnow, done_tickets, max_done, num_done);
    busy_ms += now_ms() - ts;
    return BNX_OK;
  }
};

} // namespace bnx

using namespace bnx;

struct bnx_server_s : Server {};

extern "C" {

int bnx_server_create(bnx_device d, const bnx_server_config *cfg, bnx_server *out)
{
  GUARD_BEGIN
  Device *dev = (Device *)d;
  BNX_CHECK(dev && out, BNX_ERR_INVALID, "bnx_server_create: null argument");
  bnx_server_config c{};
  if (cfg) c = *cfg;
  // default queue depth: the device's paths per batch of bnx_render (its
  // tuning), so that both forms keep as many paths in flight
  const long long paths = c.max_paths_in_flight > 0 ? c.max_paths_in_flight : dev->render_paths;
  const int act = c.max_active > 0 ? c.max_active : 256;
  BNX_CHECK(act <= 4096, BNX_ERR_INVALID, "bnx_server_create: max_active above 4096");
  BNX_CHECK(paths <= (1LL << 28), BNX_ERR_INVALID, "bnx_server_create: max_paths_in_flight above 2^28");
  BNX_CHECK(c.deadline_slack <= 1e6f && c.deadline_slack >= -1e6f, BNX_ERR_INVALID,
            "bnx_server_create: deadline_slack not a number in [-1e6, 1e6]");
  bnx_server s = new bnx_server_s;
  s->dev = dev;
  s->max_paths = paths;
  s->budget = c.frame_memory_budget > 0 ? c.frame_memory_budget : ((size_t)16 << 30);
  s->max_active = act;
  // earliest deadline first with slack 2 by default: at 0.7 load on a mixed
  // stream it keeps the p99 latency near the largest requests' time alone,
  // where shortest-remaining-first lets them starve behind medium ones
  s->slack = c.deadline_slack == 0.f ? 2.f : std::max(0.f, c.deadline_slack);
  // queue: output regions (2 rays per path in flight) plus new primaries
  int rc = s->F.ensure((size_t)(3 * paths), act, act, 2 * act + 1);
  if (rc != BNX_OK) {
    delete s;
    return rc;
  }
  BNX_HIP(hipMemsetAsync(s->F.Q.stats.ptr, 0, sizeof(unsigned long long) * STAT_COUNT * STAT_SLOTS, dev->stream));
  BNX_HIP(hipStreamSynchronize(dev->stream));
  for (int k = act - 1; k >= 0; k--) s->free_slots.push_back(k);
  *out = s;
  return BNX_OK;
  GUARD_END
}

int bnx_server_destroy(bnx_server s)
{
  BNX_CHECK(s, BNX_ERR_INVALID, "bnx_server_destroy: null server");
  delete s;
  return BNX_OK;
}

int bnx_server_submit(bnx_server s, const bnx_request *req, int64_t *ticket)
{
  GUARD_BEGIN
  BNX_CHECK(s && req && ticket, BNX_ERR_INVALID, "bnx_server_submit: null argument");
  World *w = (World *)req->world;
  BNX_CHECK(w && w->dev == s->dev, BNX_ERR_INVALID, "bnx_server_submit: world missing or of another device");
  BNX_CHECK(w->instances && w->instances->built, BNX_ERR_INVALID, "bnx_server_submit: world not built");
  BNX_CHECK(req->width > 0 && req->height > 0 && (long long)req->width * req->height <= (1LL << 26),
            BNX_ERR_INVALID, "bnx_server_submit: bad resolution");
  BNX_CHECK(req->spp >= 1 && req->spp <= MAX_BATCH_SAMPLES, BNX_ERR_INVALID, "bnx_server_submit: spp not in 1..4096");
  BNX_CHECK(req->camera.type >= BNX_CAMERA_PERSPECTIVE && req->camera.type <= BNX_CAMERA_OMNI, BNX_ERR_INVALID,
            "bnx_server_submit: unknown camera type");
  BNX_CHECK(Server::bytes_of(req->width, req->height) <= s->budget, BNX_ERR_INVALID,
            "bnx_server_submit: frame larger than the server's frame memory budget");
  Request *r = new Request;
  r->ticket = s->next_ticket++;
  r->world = w;
  r->cam = req->camera;
  r->rr = req->renderer;
  r->W = req->width;
  r->H = req->height;
  r->spp = req->spp;
  r->total_paths = (long long)req->width * req->height * req->spp;
  r->submit_ms = s->now_ms();
  s->reqs[r->ticket] = r;
  s->queued.push_back(r);
  s->submitted++;
  *ticket = r->ticket;
  return BNX_OK;
  GUARD_END
}

int bnx_server_step(bnx_server s, int64_t *done_tickets, int max_done, int *num_done)
{
  GUARD_BEGIN
  BNX_CHECK(s && num_done && max_done >= 0, BNX_ERR_INVALID, "bnx_server_step: bad argument");
  return s->step(done_tickets, max_done, num_done);
  GUARD_END
}

int bnx_server_step_n(bnx_server s, int generations, int64_t *done_tickets, int max_done, int *num_done)
{
  GUARD_BEGIN
  BNX_CHECK(s && num_done && max_done >= 0, BNX_ERR_INVALID, "bnx_server_step_n: bad argument");
  BNX_CHECK(generations >= 1 && generations <= MAX_DD_GENS, BNX_ERR_INVALID,
            "bnx_server_step_n: generations not in 1..16");
  return s->step_dd(generations, done_tickets, max_done, num_done);
  GUARD_END
}

int bnx_server_poll(bnx_server s, int64_t ticket, bnx_request_status *st)
{
  BNX_CHECK(s && st, BNX_ERR_INVALID, "bnx_server_poll: null argument");
  auto it = s->reqs.find(ticket);
  BNX_CHECK(it != s->reqs.end(), BNX_ERR_INVALID, "bnx_server_poll: unknown ticket");
  const Request *r = it->second;
  *st = bnx_request_status{};
  st->state = r->state;
  st->submit_ms = r->submit_ms;
  st->admit_ms = r->admit_ms;
  st->done_ms = r->done_ms;
  st->paths_started = r->next_path;
  st->paths_total = r->total_paths;
  return BNX_OK;
}

int bnx_server_read(bnx_server s, int64_t ticket, int channel, void *host)
{
  GUARD_BEGIN
  BNX_CHECK(s && host, BNX_ERR_INVALID, "bnx_server_read: null argument");
  auto it = s->reqs.find(ticket);
  BNX_CHECK(it != s->reqs.end(), BNX_ERR_INVALID, "bnx_server_read: unknown ticket");
  BNX_CHECK(it->second->state == BNX_REQ_DONE, BNX_ERR_INVALID, "bnx_server_read: request not done");
  return frame_read(it->second->frame, channel, host);
  GUARD_END
}

int bnx_server_release(bnx_server s, int64_t ticket)
{
  GUARD_BEGIN
  BNX_CHECK(s, BNX_ERR_INVALID, "bnx_server_release: null server");
  auto it = s->reqs.find(ticket);
  BNX_CHECK(it != s->reqs.end(), BNX_ERR_INVALID, "bnx_server_release: unknown ticket");
  Request *r = it->second;
  BNX_CHECK(r->state == BNX_REQ_DONE || r->state == BNX_REQ_FAILED, BNX_ERR_INVALID,
            "bnx_server_release: request not done");
  if (r->frame) s->pool.insert({{r->W, r->H}, r->frame});
  s->reqs.erase(it);
  delete r;
  s->released++;
  return BNX_OK;
  GUARD_END
}

int bnx_server_get_stats(bnx_server s, bnx_server_stats *st)
{
  GUARD_BEGIN
  BNX_CHECK(s && st, BNX_ERR_INVALID, "bnx_server_get_stats: null argument");
  unsigned long long host_st[STAT_COUNT];
  int rc = read_stats(s->dev, s->F, host_st);
  if (rc != BNX_OK) return rc;
  *st = bnx_server_stats{};
  st->submitted = s->submitted;
  st->admitted = s->admitted;
  st->done = s->done;
  st->released = s->released;
  st->rays_traced = s->rays;
  st->shadow_rays = (int64_t)host_st[STAT_SHADOW];
  st->volume_samples = (int64_t)host_st[STAT_SAMPLES];
  st->generations = s->gens;
  st->peak_queue = s->peak;
  st->active = (int32_t)s->active.size();
  st->queued = (int32_t)s->queued.size();
  st->frame_bytes = s->frame_bytes;
  st->busy_ms = s->busy_ms;
  st->trace_ms = s->tm.trace_ms;
  st->trace_launches = s->tm.trace_launches;
  st->march_ms = s->tm.march_ms;
  st->march_launches = s->tm.march_launches;
  st->march_rays = s->tm.march_rays;
  st->node_visits = (int64_t)host_st[STAT_NODES];
  st->tri_tests = (int64_t)host_st[STAT_TRIS];
  st->sphere_tests = (int64_t)host_st[STAT_SPHERES];
  st->mc_cells = (int64_t)host_st[STAT_CELLS];
  return BNX_OK;
  GUARD_END
}

int bnx_render_noalloc(bnx_world world, const bnx_camera *cam, const bnx_renderer *r, bnx_frame frame, int spp,
                       bnx_render_stats *stats)
{
  GUARD_BEGIN
  World *w = (World *)world;
  Frame *f = (Frame *)frame;
  BNX_CHECK(w && cam && r && f && spp >= 0, BNX_ERR_INVALID, "bnx_render_noalloc: bad argument");
  BNX_CHECK(f->dev == w->dev, BNX_ERR_INVALID, "bnx_render_noalloc: frame and world live on different devices");
  BNX_CHECK(w->instances && w->instances->built, BNX_ERR_STATE,
            "bnx_render_noalloc: the world's instance group is not built");
  BNX_CHECK(cam->type >= BNX_CAMERA_PERSPECTIVE && cam->type <= BNX_CAMERA_OMNI, BNX_ERR_INVALID,
            "bnx_render_noalloc: unknown camera type");
  BNX_CHECK((long long)f->w * f->h <= NOALLOC_PATHS, BNX_ERR_UNSUPPORTED,
            "bnx_render_noalloc: frame larger than 65536 pixels (the static scratch's limit)");
  Device *dev = w->dev;
  BNX_CHECK(dev->noalloc_queues, BNX_ERR_STATE, "bnx_render_noalloc: device without scratch");
  return render_frame_impl(w, *cam, *r, f, spp, stats, ((NoallocWavefront *)dev->noalloc_queues)->F,
                           NOALLOC_PATHS);
  GUARD_END
}

} // extern "C"

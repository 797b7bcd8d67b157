// bnx_bvh.hip -- binned-SAH BVH builder on gfx950.
//
// Replaces cuBQL::gpuBuilder as called by rtcore/cuda/Group.cu:232,387,536
// (leaf <= 4 for triangles and user geometry, leaf == 1 for instances).
//
// Split rule (per task = node-to-be over a contiguous slot range):
//   16 bins per axis over the task's centroid bounds, bin index
//   (int)((c - lo) * (16 / extent)) clamped to 15; cost of plane j is
//   area(L_j)*|L_j| + area(R_j)*|R_j| with half-area surface measure; the
//   lowest cost wins (ties: lower axis, lower plane); stable partition.
//   Median split (first half in current order) when no axis has a valid
//   plane or depth >= 48.  Tasks of <= max_leaf prims become leaves.
//
// Structure:
//   phase A -- level-synchronous over all tasks with > 32 prims: task
//     bounds and SAH bins are reduced with ordered-int atomics (LDS
//     privatised when a 2048-slot chunk lies in one task), one thread per
//     task picks the plane, and a global exclusive scan gives every slot
//     its stable destination.
//   phase B -- one thread per task of <= 32 prims builds its whole subtree
//     in place with the same split rule.
// Min/max reductions and integer counts are exact, so the tree equals the
// CPU specification's tree node for node (numbering aside).
#include "bnx_internal.h"

#include <hipcub/hipcub.hpp>

#include <chrono>

namespace bnx {
namespace {

constexpr int NB = 16;
constexpr int MAX_SAH_DEPTH = 48;
constexpr int SMALL = 32;
constexpr int BLOCK = 256;
constexpr int CHUNK = 2048;   // slots per block in the reduction kernels

struct Task {
  int begin, end, depth, parent_side;   // parent_side = parent node * 2 + side, or -1
};
struct SmallTask {
  int begin, end, depth, node;
  float clo[3], chi[3];
};
struct Split {
  int axis, bin, nl, node;
  float lo, scale;
  int pad0, pad1;
};
// per task: box lo/hi and centroid lo/hi as ordered ints
struct TB {
  int v[12];
};
// per big task: [axis][bin][count, lo xyz, hi xyz]
constexpr int BIN_INTS = 7;
constexpr int TASK_BIN_INTS = 3 * NB * BIN_INTS;

__device__ __forceinline__ float cent(const float4 &lo, const float4 &hi, int a)
{
  float l = a == 0 ? lo.x : (a == 1 ? lo.y : lo.z);
  float h = a == 0 ? hi.x : (a == 1 ? hi.y : hi.z);
  return (l + h) * 0.5f;
}
__device__ __forceinline__ int bin_of(float c, float lo, float scale)
{
  int b = (int)((c - lo) * scale);
  return b > NB - 1 ? NB - 1 : b;
}
__device__ __forceinline__ float half_area(float lx, float ly, float lz, float hx, float hy, float hz)
{
  float dx = hx - lx, dy = hy - ly, dz = hz - lz;
  return (dx * dy + dy * dz) + dz * dx;
}

__device__ void write_child(Node *nodes, int parent_side, const float *box, int ref, int cnt)
{
  int node = parent_side >> 1, side = parent_side & 1;
  float *f = (float *)&nodes[node];
  for (int k = 0; k < 6; k++) f[side * 6 + k] = box[k];
  int *ip = (int *)&nodes[node];
  ip[12 + side] = ref;
  ip[14 + side] = cnt;
}

// ------------------------------------------------------------ kernels
__global__ void k_iota(int *order, int *slot_task, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    order[i] = i;
    slot_task[i] = 0;
  }
}

__global__ void k_init_tb(TB *tb, int num)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < num) {
    for (int k = 0; k < 3; k++) {
      tb[i].v[k] = f2o(INFINITY);
      tb[i].v[3 + k] = f2o(-INFINITY);
      tb[i].v[6 + k] = f2o(INFINITY);
      tb[i].v[9 + k] = f2o(-INFINITY);
    }
  }
}

__global__ void k_init_bins(int *bins, int num_big)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < num_big * 3 * NB) {
    int *b = bins + i * BIN_INTS;
    b[0] = 0;
    for (int k = 0; k < 3; k++) {
      b[1 + k] = f2o(INFINITY);
      b[4 + k] = f2o(-INFINITY);
    }
  }
}

// task bounds (box + centroid box) of every slot with slot_task >= 0
__global__ __launch_bounds__(BLOCK) void k_bounds(int n, const int *slot_task, const int *order, const float4 *lo,
                                                  const float4 *hi, TB *tb)
{
  __shared__ int s[12];
  __shared__ int uni;
  int b = blockIdx.x * CHUNK, e = min(n, b + CHUNK);
  if (threadIdx.x == 0) {
    int t0 = slot_task[b], t1 = slot_task[e - 1];
    uni = (t0 == t1 && t0 >= 0) ? t0 : -1;
    for (int k = 0; k < 3; k++) {
      s[k] = f2o(INFINITY); s[3 + k] = f2o(-INFINITY);
      s[6 + k] = f2o(INFINITY); s[9 + k] = f2o(-INFINITY);
    }
  }
  __syncthreads();
  int u = uni;
  for (int i = b + threadIdx.x; i < e; i += BLOCK) {
    int t = u >= 0 ? u : slot_task[i];
    if (t < 0) continue;
    int p = order[i];
    float4 l = lo[p], h = hi[p];
    float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
    if (u >= 0) {
      for (int k = 0; k < 3; k++) {
        atomicMin(&s[k], f2o(lv[k]));
        atomicMax(&s[3 + k], f2o(hv[k]));
        atomicMin(&s[6 + k], f2o(c[k]));
        atomicMax(&s[9 + k], f2o(c[k]));
      }
    } else {
      int *dst = tb[t].v;
      for (int k = 0; k < 3; k++) {
        atomicMin(&dst[k], f2o(lv[k]));
        atomicMax(&dst[3 + k], f2o(hv[k]));
        atomicMin(&dst[6 + k], f2o(c[k]));
        atomicMax(&dst[9 + k], f2o(c[k]));
      }
    }
  }
  __syncthreads();
  if (u >= 0 && threadIdx.x < 12) {
    if (threadIdx.x < 3 || (threadIdx.x >= 6 && threadIdx.x < 9)) atomicMin(&tb[u].v[threadIdx.x], s[threadIdx.x]);
    else atomicMax(&tb[u].v[threadIdx.x], s[threadIdx.x]);
  }
}

// decide leaf / small / big for every task of this level
__global__ void k_classify(int num_tasks, const Task *tasks, const TB *tb, int max_leaf, Node *nodes, int *counters,
                           int *big_map, int *big_list, int *task_node, SmallTask *small, int *max_depth)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num_tasks) return;
  Task tk = tasks[t];
  int cnt = tk.end - tk.begin;
  float box[6];
  for (int k = 0; k < 6; k++) box[k] = o2f(tb[t].v[k]);
  big_map[t] = -1;
  if (cnt <= max_leaf) {
    if (tk.parent_side >= 0) write_child(nodes, tk.parent_side, box, tk.begin, cnt);
    atomicMax(max_depth, tk.depth);
    return;
  }
  int node = tk.parent_side < 0 ? 0 : atomicAdd(&counters[0], 1);
  if (tk.parent_side >= 0) write_child(nodes, tk.parent_side, box, node, 0);
  task_node[t] = node;
  if (cnt <= SMALL) {
    int si = atomicAdd(&counters[1], 1);
    SmallTask st;
    st.begin = tk.begin;
    st.end = tk.end;
    st.depth = tk.depth;
    st.node = node;
    for (int k = 0; k < 3; k++) {
      st.clo[k] = o2f(tb[t].v[6 + k]);
      st.chi[k] = o2f(tb[t].v[9 + k]);
    }
    small[si] = st;
  } else {
    int bi = atomicAdd(&counters[2], 1);
    big_map[t] = bi;
    big_list[bi] = t;
  }
}

// SAH bins of all big tasks
__global__ __launch_bounds__(BLOCK) void k_bin(int n, const int *slot_task, const int *big_map, const int *order,
                                               const float4 *lo, const float4 *hi, const TB *tb, int *bins)
{
  __shared__ int s[TASK_BIN_INTS];
  __shared__ int uni;
  __shared__ float sclo[3], sscale[3];
  int b = blockIdx.x * CHUNK, e = min(n, b + CHUNK);
  if (threadIdx.x == 0) {
    int t0 = slot_task[b], t1 = slot_task[e - 1];
    uni = (t0 == t1 && t0 >= 0 && big_map[t0] >= 0) ? t0 : -1;
  }
  __syncthreads();
  int u = uni;
  if (u >= 0) {
    for (int i = threadIdx.x; i < 3 * NB; i += BLOCK) {
      s[i * BIN_INTS] = 0;
      for (int k = 0; k < 3; k++) {
        s[i * BIN_INTS + 1 + k] = f2o(INFINITY);
        s[i * BIN_INTS + 4 + k] = f2o(-INFINITY);
      }
    }
    if (threadIdx.x < 3) {
      float clo = o2f(tb[u].v[6 + threadIdx.x]), chi = o2f(tb[u].v[9 + threadIdx.x]);
      float ext = chi - clo;
      sclo[threadIdx.x] = clo;
      sscale[threadIdx.x] = (ext >= 1e-30f) ? (float)NB / ext : -1.f;
    }
  }
  __syncthreads();
  for (int i = b + threadIdx.x; i < e; i += BLOCK) {
    int t = u >= 0 ? u : slot_task[i];
    if (t < 0) continue;
    int bi = big_map[t];
    if (bi < 0) continue;
    int p = order[i];
    float4 l = lo[p], h = hi[p];
    int lv[3] = {f2o(l.x), f2o(l.y), f2o(l.z)}, hv[3] = {f2o(h.x), f2o(h.y), f2o(h.z)};
    for (int a = 0; a < 3; a++) {
      float clo, scale;
      if (u >= 0) {
        clo = sclo[a];
        scale = sscale[a];
      } else {
        clo = o2f(tb[t].v[6 + a]);
        float ext = o2f(tb[t].v[9 + a]) - clo;
        scale = (ext >= 1e-30f) ? (float)NB / ext : -1.f;
      }
      if (scale < 0.f) continue;
      int k = bin_of(cent(l, h, a), clo, scale);
      if (u >= 0) {
        int *dst = s + (a * NB + k) * BIN_INTS;
        atomicAdd(&dst[0], 1);
        for (int q = 0; q < 3; q++) {
          atomicMin(&dst[1 + q], lv[q]);
          atomicMax(&dst[4 + q], hv[q]);
        }
      } else {
        int *dst = bins + (size_t)bi * TASK_BIN_INTS + (a * NB + k) * BIN_INTS;
        atomicAdd(&dst[0], 1);
        for (int q = 0; q < 3; q++) {
          atomicMin(&dst[1 + q], lv[q]);
          atomicMax(&dst[4 + q], hv[q]);
        }
      }
    }
  }
  __syncthreads();
  if (u >= 0) {
    int *g = bins + (size_t)big_map[u] * TASK_BIN_INTS;
    for (int i = threadIdx.x; i < 3 * NB; i += BLOCK) {
      if (s[i * BIN_INTS] == 0) continue;
      atomicAdd(&g[i * BIN_INTS], s[i * BIN_INTS]);
      for (int q = 0; q < 3; q++) {
        atomicMin(&g[i * BIN_INTS + 1 + q], s[i * BIN_INTS + 1 + q]);
        atomicMax(&g[i * BIN_INTS + 4 + q], s[i * BIN_INTS + 4 + q]);
      }
    }
  }
}

// SAH sweep over one axis' bins (identical arithmetic to the CPU spec)
__device__ void sah_axis(const float (*bl)[3], const float (*bh)[3], const int *bc, int axis, float &best_cost,
                         int &best_axis, int &best_bin)
{
  float areaL[NB], areaR[NB];
  int cntL[NB], cntR[NB];
  float al[3] = {INFINITY, INFINITY, INFINITY}, ah[3] = {-INFINITY, -INFINITY, -INFINITY};
  int cn = 0;
  for (int j = 0; j < NB - 1; j++) {
    for (int q = 0; q < 3; q++) {
      al[q] = fmn(al[q], bl[j][q]);
      ah[q] = fmx(ah[q], bh[j][q]);
    }
    cn += bc[j];
    areaL[j] = half_area(al[0], al[1], al[2], ah[0], ah[1], ah[2]);
    cntL[j] = cn;
  }
  for (int q = 0; q < 3; q++) { al[q] = INFINITY; ah[q] = -INFINITY; }
  cn = 0;
  for (int j = NB - 1; j > 0; j--) {
    for (int q = 0; q < 3; q++) {
      al[q] = fmn(al[q], bl[j][q]);
      ah[q] = fmx(ah[q], bh[j][q]);
    }
    cn += bc[j];
    areaR[j - 1] = half_area(al[0], al[1], al[2], ah[0], ah[1], ah[2]);
    cntR[j - 1] = cn;
  }
  for (int j = 0; j < NB - 1; j++) {
    if (cntL[j] == 0 || cntR[j] == 0) continue;
    float cost = areaL[j] * (float)cntL[j] + areaR[j] * (float)cntR[j];
    if (cost < best_cost) {
      best_cost = cost;
      best_axis = axis;
      best_bin = j;
    }
  }
}

// plane choice for every big task; writes the two child tasks
__global__ void k_split(int num_big, const int *big_list, const Task *tasks, const TB *tb, const int *bins,
                        const int *task_node, Split *split, Task *next)
{
  int bi = blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= num_big) return;
  int t = big_list[bi];
  Task tk = tasks[t];
  int cnt = tk.end - tk.begin;
  float best_cost = INFINITY;
  int best_axis = -1, best_bin = -1;
  const int *tbins = bins + (size_t)bi * TASK_BIN_INTS;
  if (tk.depth < MAX_SAH_DEPTH) {
    for (int a = 0; a < 3; a++) {
      float clo = o2f(tb[t].v[6 + a]);
      float ext = o2f(tb[t].v[9 + a]) - clo;
      if (!(ext >= 1e-30f)) continue;
      float bl[NB][3], bh[NB][3];
      int bc[NB];
      for (int j = 0; j < NB; j++) {
        const int *q = tbins + (a * NB + j) * BIN_INTS;
        bc[j] = q[0];
        for (int k = 0; k < 3; k++) {
          bl[j][k] = o2f(q[1 + k]);
          bh[j][k] = o2f(q[4 + k]);
        }
      }
      sah_axis(bl, bh, bc, a, best_cost, best_axis, best_bin);
    }
  }
  Split sp;
  sp.node = task_node[t];
  if (best_axis < 0) {
    sp.axis = -1;
    sp.bin = 0;
    sp.nl = cnt / 2;
    sp.lo = 0.f;
    sp.scale = 0.f;
  } else {
    float clo = o2f(tb[t].v[6 + best_axis]);
    float ext = o2f(tb[t].v[9 + best_axis]) - clo;
    sp.axis = best_axis;
    sp.bin = best_bin;
    sp.lo = clo;
    sp.scale = (float)NB / ext;
    int nl = 0;
    for (int j = 0; j <= best_bin; j++) nl += tbins[(best_axis * NB + j) * BIN_INTS];
    sp.nl = nl;
  }
  split[bi] = sp;
  next[2 * bi + 0] = Task{tk.begin, tk.begin + sp.nl, tk.depth + 1, sp.node * 2 + 0};
  next[2 * bi + 1] = Task{tk.begin + sp.nl, tk.end, tk.depth + 1, sp.node * 2 + 1};
}

__global__ void k_flags(int n, const int *slot_task, const int *big_map, const Task *tasks, const Split *split,
                        const int *order, const float4 *lo, const float4 *hi, int *flags)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = slot_task[i];
  int f = 0;
  if (t >= 0) {
    int bi = big_map[t];
    if (bi >= 0) {
      Split sp = split[bi];
      if (sp.axis < 0) {
        f = (i - tasks[t].begin) < sp.nl ? 1 : 0;
      } else {
        int p = order[i];
        f = bin_of(cent(lo[p], hi[p], sp.axis), sp.lo, sp.scale) <= sp.bin ? 1 : 0;
      }
    }
  }
  flags[i] = f;
}

__global__ void k_scatter(int n, const int *slot_task, const int *big_map, const Task *tasks, const Split *split,
                          const int *flags, const int *scan, const int *order, int *order_out, int *slot_task_out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = slot_task[i];
  int bi = t >= 0 ? big_map[t] : -1;
  if (bi < 0) {
    order_out[i] = order[i];
    slot_task_out[i] = -1;
    return;
  }
  int b = tasks[t].begin;
  int nl = split[bi].nl;
  int lrank = scan[i] - scan[b];
  int pos, child;
  if (flags[i]) {
    pos = b + lrank;
    child = 2 * bi;
  } else {
    pos = b + nl + ((i - b) - lrank);
    child = 2 * bi + 1;
  }
  order_out[pos] = order[i];
  slot_task_out[pos] = child;
}

// ---------------------------------------------------------------- phase B
__device__ void range_bounds(const int *order, int b, int e, const float4 *lo, const float4 *hi, float *box,
                             float *cb)
{
  for (int k = 0; k < 3; k++) {
    box[k] = INFINITY; box[3 + k] = -INFINITY;
    cb[k] = INFINITY; cb[3 + k] = -INFINITY;
  }
  for (int i = b; i < e; i++) {
    int p = order[i];
    float4 l = lo[p], h = hi[p];
    float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
    for (int k = 0; k < 3; k++) {
      float c = (lv[k] + hv[k]) * 0.5f;
      box[k] = fmn(box[k], lv[k]);
      box[3 + k] = fmx(box[3 + k], hv[k]);
      cb[k] = fmn(cb[k], c);
      cb[3 + k] = fmx(cb[3 + k], c);
    }
  }
}

struct SubTask {
  int begin, end, depth, node;
  float cb[6];
};

__global__ void k_small(int num, const SmallTask *tasks, int *order, const float4 *lo, const float4 *hi, Node *nodes,
                        int *counters, int max_leaf, int *max_depth)
{
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= num) return;
  SmallTask st = tasks[ti];
  SubTask stack[SMALL + 2];
  int sp = 0;
  stack[sp].begin = st.begin;
  stack[sp].end = st.end;
  stack[sp].depth = st.depth;
  stack[sp].node = st.node;
  for (int k = 0; k < 3; k++) {
    stack[sp].cb[k] = st.clo[k];
    stack[sp].cb[3 + k] = st.chi[k];
  }
  sp++;
  int tmp[SMALL];
  int local_max = 0;
  while (sp > 0) {
    SubTask tk = stack[--sp];
    int b = tk.begin, e = tk.end, n = e - b;
    float best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    if (tk.depth < MAX_SAH_DEPTH) {
      for (int a = 0; a < 3; a++) {
        float clo = tk.cb[a];
        float ext = tk.cb[3 + a] - clo;
        if (!(ext >= 1e-30f)) continue;
        float scale = (float)NB / ext;
        float bl[NB][3], bh[NB][3];
        int bc[NB];
        for (int j = 0; j < NB; j++) {
          bc[j] = 0;
          for (int k = 0; k < 3; k++) { bl[j][k] = INFINITY; bh[j][k] = -INFINITY; }
        }
        for (int i = b; i < e; i++) {
          int p = order[i];
          float4 l = lo[p], h = hi[p];
          int k = bin_of(cent(l, h, a), clo, scale);
          bc[k]++;
          bl[k][0] = fmn(bl[k][0], l.x); bl[k][1] = fmn(bl[k][1], l.y); bl[k][2] = fmn(bl[k][2], l.z);
          bh[k][0] = fmx(bh[k][0], h.x); bh[k][1] = fmx(bh[k][1], h.y); bh[k][2] = fmx(bh[k][2], h.z);
        }
        sah_axis(bl, bh, bc, a, best_cost, best_axis, best_bin);
      }
    }
    int mid;
    if (best_axis < 0) {
      mid = b + n / 2;
    } else {
      float clo = tk.cb[best_axis];
      float scale = (float)NB / (tk.cb[3 + best_axis] - clo);
      int nl = 0, nr = 0;
      for (int i = b; i < e; i++) {
        int p = order[i];
        if (bin_of(cent(lo[p], hi[p], best_axis), clo, scale) <= best_bin) order[b + nl++] = p;
        else tmp[nr++] = p;
      }
      for (int i = 0; i < nr; i++) order[b + nl + i] = tmp[i];
      mid = b + nl;
    }
    int rng[2][2] = {{b, mid}, {mid, e}};
    // push right first so the left subtree is processed first
    for (int s = 1; s >= 0; s--) {
      int cb_ = rng[s][0], ce = rng[s][1];
      float box[6], cb[6];
      range_bounds(order, cb_, ce, lo, hi, box, cb);
      int ps = tk.node * 2 + s;
      if (ce - cb_ <= max_leaf) {
        write_child(nodes, ps, box, cb_, ce - cb_);
        local_max = max(local_max, tk.depth + 1);
      } else {
        int idx = atomicAdd(&counters[0], 1);
        write_child(nodes, ps, box, idx, 0);
        SubTask c;
        c.begin = cb_;
        c.end = ce;
        c.depth = tk.depth + 1;
        c.node = idx;
        for (int k = 0; k < 6; k++) c.cb[k] = cb[k];
        stack[sp++] = c;
      }
    }
  }
  atomicMax(max_depth, local_max);
}

__global__ void k_root_leaf(Node *nodes, const TB *tb, int n)
{
  float box[6];
  for (int k = 0; k < 6; k++) box[k] = o2f(tb[0].v[k]);
  float empty[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  write_child(nodes, 0, n > 0 ? box : empty, 0, n > 0 ? n : -1);
  write_child(nodes, 1, empty, 0, -1);
}

inline int nblocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

// ------------------------------------------------------- 4-wide collapse
// One thread per 4-wide node of the current level: start from the two
// children of binary node b, then repeatedly replace the inner child of
// largest half area by its two children, in place, until there are four
// children or only leaves.  Inner children become nodes of the next level (index from
// counter[0], work item from counter[1]).  need: stack entries pushed on
// the way down to b; counter[2] collects the largest need at a leaf.
struct Cand {
  float lo[3], hi[3];
  int ref, cnt;
};

__device__ Cand cand_of(const Node &nd, int k)
{
  Cand c;
  if (k == 0) {
    c.lo[0] = nd.lo0x; c.lo[1] = nd.lo0y; c.lo[2] = nd.lo0z;
    c.hi[0] = nd.hi0x; c.hi[1] = nd.hi0y; c.hi[2] = nd.hi0z;
    c.ref = nd.ref0; c.cnt = nd.cnt0;
  } else {
    c.lo[0] = nd.lo1x; c.lo[1] = nd.lo1y; c.lo[2] = nd.lo1z;
    c.hi[0] = nd.hi1x; c.hi[1] = nd.hi1y; c.hi[2] = nd.hi1z;
    c.ref = nd.ref1; c.cnt = nd.cnt1;
  }
  return c;
}

__device__ float cand_area(const Cand &c)
{
  float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
  return (dx * dy + dy * dz) + dz * dx;
}

__global__ void k_collapse4(const Node *bin, const int4 *work, int nwork, Node4 *out, int4 *next, int *counter,
                            int cap4)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwork) return;
  const int b = work[i].x, j = work[i].y, need = work[i].z;
  Cand c[4];
  int m = 0;
  const Node root = bin[b];
  for (int k = 0; k < 2; k++) {
    Cand x = cand_of(root, k);
    if (x.cnt >= 0) c[m++] = x;
  }
  for (int round = 0; round < 3 && m < 4; round++) {
    int best = -1;
    float area = -1.f;
    for (int k = 0; k < m; k++)
      if (c[k].cnt == 0 && cand_area(c[k]) > area) {
        area = cand_area(c[k]);
        best = k;
      }
    if (best < 0) break;
    const Node nd = bin[c[best].ref];
    Cand a = cand_of(nd, 0), d = cand_of(nd, 1);
    if (a.cnt >= 0 && d.cnt >= 0) {
      // keep the binary tree's left-to-right order: a depth-first walk of
      // the 4-wide node in slot order meets the leaves in the same order
      for (int k = m; k > best + 1; k--) c[k] = c[k - 1];
      c[best] = a;
      c[best + 1] = d;
      m++;
    } else if (a.cnt >= 0 || d.cnt >= 0) {
      c[best] = a.cnt >= 0 ? a : d;
    } else {
      for (int k = best; k + 1 < m; k++) c[k] = c[k + 1];
      m--;
    }
  }
  float lo[3][4], hi[3][4];
  uint32_t code[4];
  const int child_need = need + (m > 0 ? m - 1 : 0);
  for (int k = 0; k < 4; k++) {
    if (k < m) {
      for (int a = 0; a < 3; a++) {
        lo[a][k] = c[k].lo[a];
        hi[a][k] = c[k].hi[a];
      }
      if (c[k].cnt > 0) {
        code[k] = enc_child(c[k].ref, c[k].cnt);
        atomicMax(counter + 2, child_need);
      } else {
        const int idx = atomicAdd(counter, 1);
        const int w = atomicAdd(counter + 1, 1);
        if (idx < cap4) next[w] = make_int4(c[k].ref, idx, child_need, 0);
        code[k] = (uint32_t)idx;
      }
    } else {
      for (int a = 0; a < 3; a++) {
        lo[a][k] = INFINITY;
        hi[a][k] = -INFINITY;
      }
      code[k] = EMPTY4;
    }
  }
  if (j >= cap4) return;
  Node4 n4;
  n4.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  n4.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  n4.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  n4.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  n4.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  n4.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  n4.child = make_uint4(code[0], code[1], code[2], code[3]);
  n4.pad = make_uint4(0, 0, 0, 0);
  out[j] = n4;
}

} // namespace

int collapse_bvh4(Device *dev, Bvh &bvh)
{
  hipStream_t st = dev->stream;
  // every 4-wide node but the root stands for a distinct inner binary node
  const int cap4 = bvh.num_nodes > 0 ? bvh.num_nodes : 1;
  BNX_HIP(bvh.nodes4.ensure(sizeof(Node4) * (size_t)cap4));
  DevMem work[2], ctr;
  BNX_HIP(work[0].ensure(sizeof(int4) * (size_t)cap4));
  BNX_HIP(work[1].ensure(sizeof(int4) * (size_t)cap4));
  BNX_HIP(ctr.ensure(sizeof(int) * 4));
  int h[4] = {1, 0, 0, 0};
  BNX_HIP(hipMemcpyAsync(ctr.ptr, h, sizeof(h), hipMemcpyHostToDevice, st));
  const int4 root = make_int4(0, 0, 0, 0);
  BNX_HIP(hipMemcpyAsync(work[0].ptr, &root, sizeof(root), hipMemcpyHostToDevice, st));
  int nwork = 1, cur = 0;
  for (int level = 0; nwork > 0; level++) {
    if (level > 64) {
      set_error("bvh collapse: too many levels");
      return BNX_ERR_STATE;
    }
    BNX_HIP(hipMemsetAsync((int *)ctr.ptr + 1, 0, sizeof(int), st));
    k_collapse4<<<nblocks(nwork, 128), 128, 0, st>>>(bvh.nodes.as<Node>(), work[cur].as<int4>(), nwork,
                                                     bvh.nodes4.as<Node4>(), work[cur ^ 1].as<int4>(),
                                                     ctr.as<int>(), cap4);
    BNX_HIP(hipMemcpyAsync(h, ctr.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
    BNX_HIP(hipStreamSynchronize(st));
    BNX_HIP(hipGetLastError());
    if (h[0] > cap4) {
      set_error("bvh collapse: node overflow");
      return BNX_ERR_STATE;
    }
    nwork = h[1];
    cur ^= 1;
  }
  bvh.num_nodes4 = h[0];
  bvh.stack_need4 = h[2];
  return BNX_OK;
}

int build_bvh(Device *dev, const float4 *d_lo, const float4 *d_hi, int n, int max_leaf, Bvh &out)
{
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = dev->stream;
  BuildScratch &S = dev->build;
  const int ncap = n > 0 ? n : 1;
  BNX_HIP(out.nodes.ensure(sizeof(Node) * (size_t)ncap));
  BNX_HIP(out.order.ensure(sizeof(int) * (size_t)ncap));
  BNX_HIP(S.order_tmp.ensure(sizeof(int) * (size_t)ncap));
  BNX_HIP(S.slot_task.ensure(2 * sizeof(int) * (size_t)ncap));
  BNX_HIP(S.flags.ensure(sizeof(int) * (size_t)ncap));
  BNX_HIP(S.scan.ensure(sizeof(int) * (size_t)ncap));
  BNX_HIP(S.counters.ensure(sizeof(int) * 8));
  const int max_tasks = ncap / (SMALL / 2) + 2;           // tasks with > SMALL prims -> children
  BNX_HIP(S.tasks.ensure(sizeof(Task) * (size_t)max_tasks * 2));
  BNX_HIP(S.tasks_next.ensure(sizeof(Task) * (size_t)max_tasks * 2));
  BNX_HIP(S.bounds.ensure(sizeof(TB) * (size_t)max_tasks * 2));
  BNX_HIP(S.big_map.ensure(sizeof(int) * (size_t)max_tasks * 4));
  BNX_HIP(S.bins.ensure(sizeof(int) * TASK_BIN_INTS * (size_t)max_tasks));
  BNX_HIP(S.split.ensure(sizeof(Split) * (size_t)max_tasks));
  BNX_HIP(S.small_tasks.ensure(sizeof(SmallTask) * (size_t)(ncap / 2 + 2)));

  int *order = out.order.as<int>();
  int *order_tmp = S.order_tmp.as<int>();
  int *slot_task = S.slot_task.as<int>();
  int *slot_task2 = slot_task + ncap;
  int *counters = S.counters.as<int>();   // 0: nodes, 1: small, 2: big, 3: max depth
  Task *tasks = S.tasks.as<Task>();
  Task *next = S.tasks_next.as<Task>();
  TB *tb = S.bounds.as<TB>();
  int *big_map = S.big_map.as<int>();
  int *big_list = big_map + (size_t)max_tasks * 2;
  int *task_node = big_list + max_tasks;
  Node *nodes = out.nodes.as<Node>();

  int init_ctr[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  BNX_HIP(hipMemcpyAsync(counters, init_ctr, sizeof(init_ctr), hipMemcpyHostToDevice, st));
  Task root{0, n, 0, -1};
  BNX_HIP(hipMemcpyAsync(tasks, &root, sizeof(Task), hipMemcpyHostToDevice, st));
  if (n > 0) k_iota<<<nblocks(n, 256), 256, 0, st>>>(order, slot_task, n);

  size_t cub_bytes = 0;
  BNX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_bytes, S.flags.as<int>(), S.scan.as<int>(), ncap, st));
  BNX_HIP(S.cub_tmp.ensure(cub_bytes));

  // root bounds
  k_init_tb<<<1, 32, 0, st>>>(tb, 1);
  if (n > 0) k_bounds<<<nblocks(n, CHUNK), BLOCK, 0, st>>>(n, slot_task, order, d_lo, d_hi, tb);
  TB root_tb;
  BNX_HIP(hipMemcpyAsync(&root_tb, tb, sizeof(TB), hipMemcpyDeviceToHost, st));
  BNX_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 3; k++) {
    int a = root_tb.v[k], b = root_tb.v[3 + k];
    float fa, fb;
    int ia = a >= 0 ? a : (a ^ 0x7fffffff), ib = b >= 0 ? b : (b ^ 0x7fffffff);
    memcpy(&fa, &ia, 4);
    memcpy(&fb, &ib, 4);
    out.root_lo[k] = fa;
    out.root_hi[k] = fb;
  }

  int host_ctr[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  if (n <= max_leaf) {
    k_root_leaf<<<1, 1, 0, st>>>(nodes, tb, n);
    host_ctr[3] = n > 0 ? 1 : 0;
  } else {
    int num_tasks = 1;
    bool first = true;
    while (num_tasks > 0) {
      if (!first) {
        k_init_tb<<<nblocks(num_tasks, 256), 256, 0, st>>>(tb, num_tasks);
        k_bounds<<<nblocks(n, CHUNK), BLOCK, 0, st>>>(n, slot_task, order, d_lo, d_hi, tb);
      }
      first = false;
      BNX_HIP(hipMemsetAsync(counters + 2, 0, sizeof(int), st));
      k_classify<<<nblocks(num_tasks, 128), 128, 0, st>>>(num_tasks, tasks, tb, max_leaf, nodes, counters, big_map,
                                                          big_list, task_node, S.small_tasks.as<SmallTask>(),
                                                          counters + 3);
      int num_big = 0;
      BNX_HIP(hipMemcpyAsync(&num_big, counters + 2, sizeof(int), hipMemcpyDeviceToHost, st));
      BNX_HIP(hipStreamSynchronize(st));
      if (num_big == 0) break;
      if (2 * num_big > 2 * max_tasks) {
        set_error("bvh builder: task overflow");
        return BNX_ERR_STATE;
      }
      k_init_bins<<<nblocks((long)num_big * 3 * NB, 256), 256, 0, st>>>(S.bins.as<int>(), num_big);
      k_bin<<<nblocks(n, CHUNK), BLOCK, 0, st>>>(n, slot_task, big_map, order, d_lo, d_hi, tb, S.bins.as<int>());
      k_split<<<nblocks(num_big, 64), 64, 0, st>>>(num_big, big_list, tasks, tb, S.bins.as<int>(), task_node,
                                                   S.split.as<Split>(), next);
      k_flags<<<nblocks(n, 256), 256, 0, st>>>(n, slot_task, big_map, tasks, S.split.as<Split>(), order, d_lo, d_hi,
                                               S.flags.as<int>());
      size_t tb_bytes = cub_bytes;
      BNX_HIP(hipcub::DeviceScan::ExclusiveSum(S.cub_tmp.ptr, tb_bytes, S.flags.as<int>(), S.scan.as<int>(), n, st));
      k_scatter<<<nblocks(n, 256), 256, 0, st>>>(n, slot_task, big_map, tasks, S.split.as<Split>(), S.flags.as<int>(),
                                                 S.scan.as<int>(), order, order_tmp, slot_task2);
      BNX_HIP(hipMemcpyAsync(order, order_tmp, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
      std::swap(slot_task, slot_task2);
      std::swap(tasks, next);
      num_tasks = 2 * num_big;
    }
    BNX_HIP(hipMemcpyAsync(host_ctr, counters, sizeof(host_ctr), hipMemcpyDeviceToHost, st));
    BNX_HIP(hipStreamSynchronize(st));
    if (host_ctr[1] > 0)
      k_small<<<nblocks(host_ctr[1], 64), 64, 0, st>>>(host_ctr[1], S.small_tasks.as<SmallTask>(), order, d_lo, d_hi,
                                                       nodes, counters, max_leaf, counters + 3);
    BNX_HIP(hipMemcpyAsync(host_ctr, counters, sizeof(host_ctr), hipMemcpyDeviceToHost, st));
  }
  BNX_HIP(hipStreamSynchronize(st));
  BNX_HIP(hipGetLastError());
  out.num_nodes = host_ctr[0];
  out.num_prims = n;
  out.max_depth = host_ctr[3];
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return BNX_OK;
}

} // namespace bnx

// bnx_volume.hip -- volumes and isosurfaces on gfx950: structured and
// tetrahedral fields, macrocell value ranges, majorants, and the DDA +
// Woodcock (delta tracking) and isosurface march.
//
// Replaces MCVolumeAccel (volume/MCAccelerator.h:155-195, 550-689) and
// MCIsoSurfaceAccel (:325-548) over StructuredDataSampler and
// UMeshCuBQLSampler, StructuredData_computeMCs (volume/StructuredData.cu:
// 30-62), umeshRasterCells (umesh/common/UMeshField.cu:28-161), mapMCs
// (volume/MCGrid.cu:96-113), dda3 (volume/DDA.h:60-201) and
// Woodcock::sampleRange (volume/Volume.h:154-203).
//
// Design: isosurfaces and volumes are handled in their own wavefront kernel
// after the surface trace (one thread per queued ray), clipped to the
// closest hit so far.  The march is long, data-dependent and gather-bound,
// so keeping it out of the BVH traversal kernel keeps that kernel's
// registers and divergence low.  Structured fields are filtered in software
// on the raw float grid (hardware texture filtering uses 8-bit fixed-point
// weights); a 256^3 field is 64 MiB and stays resident in the 256 MiB
// Infinity Cache.  Tet fields keep vertices as float4 (position, scalar: one
// 16-byte load per corner) and tets as int4 in element-BVH slot order; the
// point query's stack is a 16-entry ring per thread in LDS.
#include "bnx_internal.h"
#include "bnx_render.h"
#include "bnx_volume.h"

#include <cmath>

#include <hipcub/hipcub.hpp>

namespace bnx {
namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// StructuredData_computeMCs: min/max over the (cpm+1)^3 vertices of a macrocell
__global__ void k_mc_ranges(VolumeDev v, float2 *ranges)
{
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int n = v.mc[0] * v.mc[1] * v.mc[2];
  if (tid >= n) return;
  int mx = tid % v.mc[0], my = (tid / v.mc[0]) % v.mc[1], mz = tid / (v.mc[0] * v.mc[1]);
  const int c = v.cells_per_mc;
  float lo = INFINITY, hi = -INFINITY;
  for (int iz = 0; iz <= c; iz++) {
    int sz = mz * c + iz;
    if (sz >= v.dims[2]) continue;
    for (int iy = 0; iy <= c; iy++) {
      int sy = my * c + iy;
      if (sy >= v.dims[1]) continue;
      const float *row = v.scalars + (size_t)v.dims[0] * ((size_t)sy + (size_t)v.dims[1] * sz);
      for (int ix = 0; ix <= c; ix++) {
        int sx = mx * c + ix;
        if (sx >= v.dims[0]) continue;
        float f = row[sx];
        lo = fmn(lo, f);
        hi = fmx(hi, f);
      }
    }
  }
  ranges[tid] = make_float2(lo, hi);
}

// mapMCs + TransferFunction::DD::majorant
__global__ void k_majorants(VolumeDev v, const float2 *ranges, float *maj)
{
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int n = v.mc[0] * v.mc[1] * v.mc[2];
  if (tid >= n) return;
  float2 r = ranges[tid];
  maj[tid] = tf_majorant(v, r.x, r.y);
}

// ------------------------------------------------------------------ cells
BNX_D int cell_nverts(int type) { return type == CELL_HEX ? 8 : (type == CELL_PRISM ? 6 : (type == CELL_PYR ? 5 : 4)); }

// the vertices (position, scalar) of cell c: k < cell_nverts
BNX_D float4 cell_vertex(const int4 *tets, const int4 *ext, const float4 *verts, int c, int k)
{
  const int4 q = k < 4 ? tets[c] : ext[c];
  const int j = k & 3;
  return verts[j == 0 ? q.x : (j == 1 ? q.y : (j == 2 ? q.z : q.w))];
}

// cellBounds (UMeshField.h:270-345): box of the cell's vertices, position
// in xyz and scalar range in w
BNX_D void cell_box(const int4 *tets, const int4 *ext, const uint8_t *types, const float4 *verts, int c, float4 &lo,
                    float4 &hi)
{
  const int nv = types ? cell_nverts(types[c]) : 4;
  lo = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
  hi = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
  for (int k = 0; k < nv; k++) {
    const float4 p = cell_vertex(tets, ext, verts, c, k);
    lo = make_float4(fmn(lo.x, p.x), fmn(lo.y, p.y), fmn(lo.z, p.z), fmn(lo.w, p.w));
    hi = make_float4(fmx(hi.x, p.x), fmx(hi.y, p.y), fmx(hi.z, p.z), fmx(hi.w, p.w));
  }
}

// world-space box of every cell (element-BVH build input); ext and types
// null: every cell is a tet
__global__ void k_cell_boxes(const int4 *tets, const int4 *ext, const uint8_t *types, const float4 *verts, int n,
                             float4 *lo, float4 *hi)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float4 l, h;
  cell_box(tets, ext, types, verts, t, l, h);
  lo[t] = make_float4(l.x, l.y, l.z, 0.f);
  hi[t] = make_float4(h.x, h.y, h.z, 0.f);
}

// cells in element-BVH slot order
__global__ void k_cell_sort(const int4 *tets, const int4 *ext, const uint8_t *types, const int *order, int n,
                            int4 *tets_out, int4 *ext_out, uint8_t *types_out)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int c = order[s];
  tets_out[s] = tets[c];
  if (types) {
    ext_out[s] = ext[c];
    types_out[s] = types[c];
  }
}

// the 80-byte tet records (VolumeDev::utetv) in slot order: four vertices
// and the face neighbours
__global__ void k_tet_pack(const int4 *tets, const float4 *verts, const int4 *nbr, int n, float4 *rec)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 t = tets[s];
  float4 *r = rec + 5 * (size_t)s;
  r[0] = verts[t.x];
  r[1] = verts[t.y];
  r[2] = verts[t.z];
  r[3] = verts[t.w];
  const int4 b = nbr[s];
  r[4] = make_float4(__int_as_float(b.x), __int_as_float(b.y), __int_as_float(b.z), __int_as_float(b.w));
}

__global__ void k_fill_ranges(int *lo, int *hi, int n)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  lo[c] = f2o(INFINITY);
  hi[c] = f2o(-INFINITY);
}

__global__ void k_finish_ranges(const int *lo, const int *hi, int n, float2 *ranges)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) ranges[c] = make_float2(o2f(lo[c]), o2f(hi[c]));
}

BNX_D float4 mid4(float4 a, float4 b)
{
  return make_float4(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.5f * (a.w + b.w));
}
BNX_D float len3(float4 a, float4 b) { return length(xyz(b) - xyz(a)); }

// macrocell min/max of one cell: a per-thread cache of the last cell, since
// the 32 pieces of a tet mostly fall into the same cell; flushed with
// ordered-int atomics (exact float min/max in any order)
struct RangeCache {
  int cell = -1;
  float lo = INFINITY, hi = -INFINITY;
  BNX_D void flush(int *rlo, int *rhi)
  {
    if (cell < 0) return;
    atomicMin(rlo + cell, f2o(lo));
    atomicMax(rhi + cell, f2o(hi));
  }
  BNX_D void add(int c, float l, float h, int *rlo, int *rhi)
  {
    if (c != cell) {
      flush(rlo, rhi);
      cell = c;
      lo = l;
      hi = h;
    } else {
      lo = fmn(lo, l);
      hi = fmx(hi, h);
    }
  }
};

// umeshRasterCells + rasterTet<5> (UMeshField.cu:28-161): each tet is split
// 5 times on its longest edge (ties: ab, ac, ad, bc, bd, cd in that order);
// the box of each of the 32 pieces is rasterized with rasterBox
// (volume/MCGrid.cuh:95-122), which skips boxes flat in any dimension.
// Piece p is reached by taking child (p >> (4 - level)) & 1 at every level.
// Other cells rasterize their own box (the three-argument rasterBox,
// MCGrid.cuh:66-93, which skips only inverted boxes).
BNX_D void raster_cell_box(const VolumeDev &v, float4 l, float4 h, RangeCache &cache, int *rlo, int *rhi)
{
  if (l.x > h.x || l.y > h.y || l.z > h.z) return;
  int lo[3], hi[3];
  const float bl[3] = {l.x, l.y, l.z}, bh[3] = {h.x, h.y, h.z};
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)((bl[k] - v.mc_origin[k]) * v.mc_rcp_spacing[k]);
    hi[k] = (int)((bh[k] - v.mc_origin[k]) * v.mc_rcp_spacing[k]);
    lo[k] = min(max(lo[k], 0), v.mc[k] - 1);
    hi[k] = min(max(hi[k], 0), v.mc[k] - 1);
  }
  for (int iz = lo[2]; iz <= hi[2]; iz++)
    for (int iy = lo[1]; iy <= hi[1]; iy++)
      for (int ix = lo[0]; ix <= hi[0]; ix++) cache.add(ix + v.mc[0] * (iy + v.mc[1] * iz), l.w, h.w, rlo, rhi);
}

__global__ void k_umesh_raster(const int4 *tets, const int4 *ext, const uint8_t *types, const float4 *verts, int n,
                               VolumeDev v, int *rlo, int *rhi)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (types && types[t] != CELL_TET) {
    RangeCache cache;
    float4 l, h;
    cell_box(tets, ext, types, verts, t, l, h);
    raster_cell_box(v, l, h, cache, rlo, rhi);
    cache.flush(rlo, rhi);
    return;
  }
  const int4 q = tets[t];
  const float4 A = verts[q.x], B = verts[q.y], Cc = verts[q.z], D = verts[q.w];
  RangeCache cache;
  for (int p = 0; p < 32; p++) {
    float4 a = A, b = B, c = Cc, d = D;
    for (int level = 0; level < 5; level++) {
      const float lab = len3(a, b), lac = len3(a, c), lad = len3(a, d);
      const float lbc = len3(b, c), lbd = len3(b, d), lcd = len3(c, d);
      const float maxLen = fmx(fmx(fmx(fmx(fmx(lab, lac), lad), lbc), lbd), lcd);
      float4 oa, ob, oc, od0, od1;
      if (lab >= maxLen) { oa = mid4(a, b); ob = c; oc = d; od0 = a; od1 = b; }
      else if (lac >= maxLen) { oa = mid4(a, c); ob = b; oc = d; od0 = a; od1 = c; }
      else if (lad >= maxLen) { oa = mid4(a, d); ob = b; oc = c; od0 = a; od1 = d; }
      else if (lbc >= maxLen) { oa = mid4(b, c); ob = a; oc = d; od0 = b; od1 = c; }
      else if (lbd >= maxLen) { oa = mid4(b, d); ob = a; oc = c; od0 = b; od1 = d; }
      else { oa = mid4(c, d); ob = a; oc = b; od0 = c; od1 = d; }
      a = oa;
      b = ob;
      c = oc;
      d = ((p >> (4 - level)) & 1) ? od1 : od0;
    }
    const float lx = fmn(fmn(fmn(a.x, b.x), c.x), d.x), hx = fmx(fmx(fmx(a.x, b.x), c.x), d.x);
    const float ly = fmn(fmn(fmn(a.y, b.y), c.y), d.y), hy = fmx(fmx(fmx(a.y, b.y), c.y), d.y);
    const float lz = fmn(fmn(fmn(a.z, b.z), c.z), d.z), hz = fmx(fmx(fmx(a.z, b.z), c.z), d.z);
    const float lw = fmn(fmn(fmn(a.w, b.w), c.w), d.w), hw = fmx(fmx(fmx(a.w, b.w), c.w), d.w);
    if (lx >= hx || ly >= hy || lz >= hz) continue;
    int lo[3], hi[3];
    const float bl[3] = {lx, ly, lz}, bh[3] = {hx, hy, hz};
    for (int k = 0; k < 3; k++) {
      lo[k] = (int)((bl[k] - v.mc_origin[k]) * v.mc_rcp_spacing[k]);
      hi[k] = (int)((bh[k] - v.mc_origin[k]) * v.mc_rcp_spacing[k]);
      lo[k] = min(max(lo[k], 0), v.mc[k] - 1);
      hi[k] = min(max(hi[k], 0), v.mc[k] - 1);
    }
    for (int iz = lo[2]; iz <= hi[2]; iz++)
      for (int iy = lo[1]; iy <= hi[1]; iy++)
        for (int ix = lo[0]; ix <= hi[0]; ix++)
          cache.add(ix + v.mc[0] * (iy + v.mc[1] * iz), lw, hw, rlo, rhi);
  }
  cache.flush(rlo, rhi);
}

// ----------------------------------------------------------------- march
// March every ray of the generation's segments through its request's
// isosurfaces (first) and volumes, one ray per thread, each clipped to the
// closest hit so far; each block lies inside one request's segment, so the
// tables are block-uniform.  With no surface instances in the batch the
// trace kernel is skipped and every ray enters here as a surface miss
// (surfaces == 0).  MODE 0: every volume of the batch is structured and
// there are no isosurfaces; MARCH_FULL: isosurfaces and unstructured fields
// may occur; with MARCH_CELLS, fields with other cells than tets too (a
// kernel of its own: the cells' Newton inversions would otherwise raise the
// register pressure of the tet kernel).
// waves per SIMD the march kernel is compiled for, lean and full variant.
// The full variant (tet point location + isosurfaces) needs 126 VGPRs to
// run without spilling, which allows 4 waves per SIMD.  Compiled for 6 (80
// VGPRs) it spills 112 of them to scratch, and the spill stores then make up
// most of its memory writes (4.9 GB per launch against 0.95 GB at 4 waves),
// yet on the 20.5M-tet volume with an isosurface (1024^2, 32 spp) a launch
// takes 24.2 ms against 27.9 ms at 4 waves and 25.2 ms at 5 (40 spilled):
// the dependent gathers of the point queries need the extra waves to hide
// their latency more than the spills cost.  Splitting the kernel into an
// isosurface pass and a volume pass would not remove the spills (each alone
// spills 46 and 64 VGPRs at 6 waves).
#ifndef BNX_MARCH_WAVES_LEAN
#define BNX_MARCH_WAVES_LEAN 5
#endif
#ifndef BNX_MARCH_WAVES_FULL
#define BNX_MARCH_WAVES_FULL 6
#endif
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE ? BNX_MARCH_WAVES_FULL
                                                                          : BNX_MARCH_WAVES_LEAN))) void
k_volume_march(const SlotDev *__restrict__ slots, SegView sv, RayQueue q,
                                                      int surfaces, unsigned long long *stats)
{
  constexpr bool FULL = MODE != 0, CELLS = (MODE & MARCH_CELLS) != 0;
  __shared__ uint32_t ring[FULL ? PQ_SHORT * PQ_TB : 1];
  MarchCtx cx{ring + threadIdx.x, 0, 0, 0, 0};
  int i;
  bool valid;
  const Segment sg = segment_of<256>(sv, i, valid);
  const SlotDev &SL = slots[sg.slot];
  const VolumeList vl{SL.vols, SL.num_vols};
  const IsoDev BNX_G *isos = SL.isos;
  const int niso = FULL ? SL.num_isos : 0;
  if (valid) {
    float4 o = q.org[i];
    float4 d = q.dir[i];
    int4 hit = surfaces ? q.hit[i] : make_int4(-1, -1, -1, 0);
    const bool shadow = (__float_as_uint(d.w) & FLAG_SHADOW) != 0;
    if (!(shadow && hit.x >= 0)) {                 // an occluded shadow ray is done
      const v3 org = xyz(o), dir = fix_dir(xyz(d));
      float tcur = o.w;
      uint64_t rng = q.rng[i];
      const uint64_t rng0 = rng;
      int ik = -1;
      float it = 0.f;
      v3 iN = splat(0.f);
      for (int k = 0; k < niso; k++) {
        float t;
        v3 N;
        if (iso_march<CELLS>(isos[k], org, dir, 0.f, tcur, rng, t, N, cx)) {
          ik = k;
          it = t;
          iN = N;
          tcur = t;
          if (shadow) break;
        }
      }
      int vk = -1;
      float vt = 0.f, rgb[3] = {0.f, 0.f, 0.f};
      // ratio tracking: a shadow ray's transmittance estimate through the
      // volumes scales its MIS weight (the light sample it carries); zero
      // transmittance occludes it as a delta-tracking event does
      const bool ratio = shadow && (SL.rr.flags & BNX_RENDER_RATIO_TRACKING) != 0;
      float T = 1.f;
      for (int k = 0; k < vl.count && !(shadow && ik >= 0); k++) {
        float t, c[3];
        if (ratio) {
          T *= ratio_track<FULL, CELLS>(vl.vols[k], org, dir, 0.f, tcur, rng, cx);
          if (T == 0.f) {
            vk = k;
            vt = tcur;
            break;
          }
          continue;
        }
        if (volume_march<FULL, CELLS>(vl.vols[k], org, dir, 0.f, tcur, rng, t, c, cx)) {
          vk = k;
          vt = t;
          tcur = t;
          rgb[0] = c[0]; rgb[1] = c[1]; rgb[2] = c[2];
          if (shadow) break;
        }
      }
      if (rng != rng0) q.rng[i] = rng;
      if (ratio && vk < 0 && T < 1.f) q.ps[i].w = __float_as_uint(__uint_as_float(q.ps[i].w) * T);
      if (vk >= 0) {
        q.org[i].w = vt;
        hit = make_int4(0, VOLUME_GEOM_BASE + vk, vl.vols[vk].rtc_inst, 0);
        q.vcol[i] = make_float4(rgb[0], rgb[1], rgb[2], 0.f);
        q.hit[i] = hit;
      } else if (FULL && ik >= 0) {
        q.org[i].w = it;
        hit = make_int4(0, ISO_GEOM_BASE + ik, isos[ik].rtc_inst, 0);
        q.vcol[i] = make_float4(iN.x, iN.y, iN.z, 0.f);
        q.hit[i] = hit;
      } else if (!surfaces) {
        q.hit[i] = hit;
      }
    } else if (!surfaces) {
      q.hit[i] = hit;
    }
  }
  const uint32_t s = wave_sum(cx.samples), c = wave_sum(cx.cells);
  const uint32_t un = wave_sum(cx.unodes), ut = wave_sum(cx.utets);
  if (lane_id() == 0) {
    stat_add(stats, STAT_SAMPLES, s);
    stat_add(stats, STAT_CELLS, c);
    stat_add(stats, STAT_UNODES, un);
    stat_add(stats, STAT_UTETS, ut);
  }
}

// one thread per brick texel: the texel of the linear field at the brick's
// cell origin + (lx, ly, lz), clamped to the field (apron texels past the
// last vertex repeat it; no footprint reads them)
__global__ void k_brick(const float *src, int nx, int ny, int nz, int shift, int nbx, int nby, long long total,
                        float *dst)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int E = (1 << shift) + 1, E3 = E * E * E;
  const long long b = i / E3;
  const int t = (int)(i - b * E3);
  const int lx = t % E, ly = (t / E) % E, lz = t / (E * E);
  const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long long)nbx * nby));
  const int x = min((bx << shift) + lx, nx - 1), y = min((by << shift) + ly, ny - 1),
            z = min((bz << shift) + lz, nz - 1);
  dst[i] = src[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)];
}

} // namespace

int volume_brick(Volume *v, int shift)
{
  VolumeDev &d = v->dd;
  if (shift <= 0) {
    v->bricks.release();
    d.bricks = nullptr;
    d.brick_shift = 0;
    return BNX_OK;
  }
  const int B = 1 << shift, E = B + 1;
  for (int k = 0; k < 3; k++) d.nbricks[k] = (d.ncells[k] + B - 1) / B;
  const long long total = (long long)d.nbricks[0] * d.nbricks[1] * d.nbricks[2] * E * E * E;
  BNX_CHECK(total < (1LL << 31), BNX_ERR_UNSUPPORTED, "structured field too large for the bricked layout");
  BNX_HIP(v->bricks.ensure(sizeof(float) * (size_t)total));
  k_brick<<<nblk(total, 256), 256, 0, v->dev->stream>>>(v->scalars.as<float>(), d.dims[0], d.dims[1], d.dims[2],
                                                        shift, d.nbricks[0], d.nbricks[1], total,
                                                        v->bricks.as<float>());
  BNX_HIP(hipGetLastError());
  BNX_SET(d.bricks, v->bricks.ptr);
  d.brick_shift = shift;
  v->version++;
  return BNX_OK;
}

int volume_build(Volume *v)
{
  Device *dev = v->dev;
  VolumeDev &d = v->dd;
  const size_t nmc = (size_t)d.mc[0] * d.mc[1] * d.mc[2];
  BNX_HIP(v->ranges.ensure(sizeof(float2) * nmc));
  BNX_HIP(v->majorants.ensure(sizeof(float) * nmc));
  BNX_SET(d.majorants, v->majorants.ptr);
  BNX_SET(d.ranges, v->ranges.ptr);
  k_mc_ranges<<<nblk((long)nmc, 64), 64, 0, dev->stream>>>(d, v->ranges.as<float2>());
  k_majorants<<<nblk((long)nmc, 256), 256, 0, dev->stream>>>(d, v->ranges.as<float2>(), v->majorants.as<float>());
  BNX_HIP(hipStreamSynchronize(dev->stream));
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

// ------------------------------------------------- tet face adjacency
// Face k of a tet is the one opposite its vertex k.  Every face gets a
// 64-bit hash of its sorted vertex triple; after sorting the hashes, a run of
// exactly two faces with the same triple is a shared face.  Hash collisions
// and non-manifold faces leave the neighbour at -1 (the walk then falls back
// to the depth-first query), so the table only ever names true neighbours.
__device__ uint64_t mix64(uint64_t x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__device__ int3 face_of(const int4 t, int k)
{
  int a, b, c;
  if (k == 0) { a = t.y; b = t.z; c = t.w; }
  else if (k == 1) { a = t.x; b = t.z; c = t.w; }
  else if (k == 2) { a = t.x; b = t.y; c = t.w; }
  else { a = t.x; b = t.y; c = t.z; }
  // sort the triple
  if (a > b) { int x = a; a = b; b = x; }
  if (b > c) { int x = b; b = c; c = x; }
  if (a > b) { int x = a; a = b; b = x; }
  return make_int3(a, b, c);
}

__global__ void k_face_keys(const int4 *tets, int ntet, uint64_t *keys, int *vals, int4 *nbr)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntet) return;
  const int4 t = tets[i];
  for (int k = 0; k < 4; k++) {
    const int3 f = face_of(t, k);
    keys[4 * (size_t)i + k] =
        mix64(((uint64_t)(uint32_t)f.x << 32 | (uint32_t)f.y) ^ mix64((uint64_t)(uint32_t)f.z + 0x9e3779b97f4a7c15ull));
    vals[4 * (size_t)i + k] = 4 * i + k;
  }
  nbr[i] = make_int4(-1, -1, -1, -1);
}

__global__ void k_face_match(const int4 *tets, const uint64_t *keys, const int *vals, long long nf, int4 *nbr)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= nf) return;
  if (keys[i] != keys[i + 1]) return;
  if (i > 0 && keys[i - 1] == keys[i]) return;         // not the first of its run
  if (i + 2 < nf && keys[i + 2] == keys[i]) return;    // a run of more than two
  const int fa = vals[i], fb = vals[i + 1];
  const int ta = fa >> 2, ka = fa & 3, tb = fb >> 2, kb = fb & 3;
  const int3 A = face_of(tets[ta], ka), B = face_of(tets[tb], kb);
  if (A.x != B.x || A.y != B.y || A.z != B.z || ta == tb) return;
  ((int *)nbr)[4 * (size_t)ta + ka] = tb;
  ((int *)nbr)[4 * (size_t)tb + kb] = ta;
}

int build_tet_neighbors(Device *dev, const int4 *tets, int ntet, DevMem &nbr)
{
  hipStream_t st = dev->stream;
  const size_t nf = 4 * (size_t)ntet;
  DevMem keys, keys2, vals, vals2, tmp;
  BNX_HIP(nbr.ensure(sizeof(int4) * (size_t)(ntet > 0 ? ntet : 1)));
  if (ntet == 0) return BNX_OK;
  BNX_HIP(keys.ensure(sizeof(uint64_t) * nf));
  BNX_HIP(keys2.ensure(sizeof(uint64_t) * nf));
  BNX_HIP(vals.ensure(sizeof(int) * nf));
  BNX_HIP(vals2.ensure(sizeof(int) * nf));
  k_face_keys<<<nblk(ntet, 256), 256, 0, st>>>(tets, ntet, keys.as<uint64_t>(), vals.as<int>(), nbr.as<int4>());
  size_t tmp_bytes = 0;
  BNX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             vals.as<int>(), vals2.as<int>(), (int)nf, 0, 64, st));
  BNX_HIP(tmp.ensure(tmp_bytes));
  BNX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             vals.as<int>(), vals2.as<int>(), (int)nf, 0, 64, st));
  k_face_match<<<nblk((long)nf, 256), 256, 0, st>>>(tets, keys2.as<uint64_t>(), vals2.as<int>(), (long long)nf,
                                                   nbr.as<int4>());
  BNX_HIP(hipGetLastError());
  BNX_HIP(hipStreamSynchronize(st));
  return BNX_OK;
}

// the seed grid: per cell, the tet (depth-first query) holding its centre
__global__ __launch_bounds__(PQ_TB) void k_umesh_seeds(VolumeDev d, int ncells, int *seeds)
{
  __shared__ uint32_t ring[PQ_SHORT * PQ_TB];
  MarchCtx cx{ring + threadIdx.x, 0, 0, 0, 0};
  const int c = blockIdx.x * PQ_TB + threadIdx.x;
  if (c >= ncells) return;
  const int nx = d.seed_dims[0], ny = d.seed_dims[1];
  const int ix = c % nx, iy = (c / nx) % ny, iz = c / (nx * ny);
  const v3 P = v3{d.seed_origin[0] + ((float)ix + .5f) / d.seed_rcp[0],
                  d.seed_origin[1] + ((float)iy + .5f) / d.seed_rcp[1],
                  d.seed_origin[2] + ((float)iz + .5f) / d.seed_rcp[2]};
  float f;
  seeds[c] = umesh_dfs(d, P, cx, f);
}

// seed grid of about one cell per four tets over the bounds (at most 256 a
// side); an empty axis gets one cell
int build_seed_grid(Volume *v, int ntet)
{
  VolumeDev &d = v->dd;
  double ext[3], vol = 1.0;
  int flat = 0;
  for (int k = 0; k < 3; k++) {
    ext[k] = (double)d.hi[k] - (double)d.lo[k];
    if (ext[k] > 0.0) vol *= ext[k];
    else flat++;
  }
  const double h = std::cbrt(vol / std::max(1.0, ntet / 4.0));
  long long ncells = 1;
  for (int k = 0; k < 3; k++) {
    int n = 1;
    if (ext[k] > 0.0 && flat == 0) n = (int)std::min(256.0, std::max(1.0, std::ceil(ext[k] / h)));
    d.seed_dims[k] = n;
    d.seed_origin[k] = d.lo[k];
    d.seed_rcp[k] = ext[k] > 0.0 ? (float)n / (float)ext[k] : 0.f;
    ncells *= n;
  }
  if (flat > 0) {
    // a flat mesh holds no tet with volume: no seeds
    d.useed = nullptr;
    return BNX_OK;
  }
  BNX_HIP(v->useed.ensure(sizeof(int) * (size_t)ncells));
  k_umesh_seeds<<<nblk(ncells, PQ_TB), PQ_TB, 0, v->dev->stream>>>(d, (int)ncells, v->useed.as<int>());
  BNX_HIP(hipGetLastError());
  BNX_SET(d.useed, v->useed.ptr);
  return BNX_OK;
}

int volume_build_umesh(Volume *v, const int4 *tets4, const int4 *ext4, const uint8_t *types, int ntet,
                       const float4 *verts4, int nvert)
{
  Device *dev = v->dev;
  VolumeDev &d = v->dd;
  hipStream_t st = dev->stream;
  const bool mixed = types != nullptr;
  BNX_HIP(v->uverts.ensure(sizeof(float4) * (size_t)nvert));
  BNX_HIP(hipMemcpyAsync(v->uverts.ptr, verts4, sizeof(float4) * (size_t)nvert, hipMemcpyHostToDevice, st));
  DevMem tets, ext, typ, lo, hi, rlo, rhi;
  BNX_HIP(tets.ensure(sizeof(int4) * (size_t)ntet));
  BNX_HIP(hipMemcpyAsync(tets.ptr, tets4, sizeof(int4) * (size_t)ntet, hipMemcpyHostToDevice, st));
  if (mixed) {
    BNX_HIP(ext.ensure(sizeof(int4) * (size_t)ntet));
    BNX_HIP(typ.ensure((size_t)ntet));
    BNX_HIP(hipMemcpyAsync(ext.ptr, ext4, sizeof(int4) * (size_t)ntet, hipMemcpyHostToDevice, st));
    BNX_HIP(hipMemcpyAsync(typ.ptr, types, (size_t)ntet, hipMemcpyHostToDevice, st));
  }
  const int4 *ext_d = mixed ? ext.as<int4>() : nullptr;
  const uint8_t *typ_d = mixed ? typ.as<uint8_t>() : nullptr;
  // element BVH over the cells' boxes; cells stored in its slot order
  BNX_HIP(lo.ensure(sizeof(float4) * (size_t)ntet));
  BNX_HIP(hi.ensure(sizeof(float4) * (size_t)ntet));
  k_cell_boxes<<<nblk(ntet, 256), 256, 0, st>>>(tets.as<int4>(), ext_d, typ_d, v->uverts.as<float4>(), ntet,
                                                lo.as<float4>(), hi.as<float4>());
  BNX_HIP(hipGetLastError());
  int rc = build_bvh(dev, lo.as<float4>(), hi.as<float4>(), ntet, 4, v->ebvh);
  if (rc != BNX_OK) return rc;
  if (v->ebvh.max_depth + 1 > PQ_SHORT + PQ_DEEP) {
    set_error("bnx_volume_create_umesh: element BVH deeper than the point-query stack");
    return BNX_ERR_UNSUPPORTED;
  }
  BNX_HIP(v->utets.ensure(sizeof(int4) * (size_t)ntet));
  if (mixed) {
    BNX_HIP(v->uext.ensure(sizeof(int4) * (size_t)ntet));
    BNX_HIP(v->utypes.ensure((size_t)ntet));
  } else {
    v->uext.release();
    v->utypes.release();
  }
  k_cell_sort<<<nblk(ntet, 256), 256, 0, st>>>(tets.as<int4>(), ext_d, typ_d, v->ebvh.order.as<int>(), ntet,
                                               v->utets.as<int4>(), v->uext.as<int4>(), v->utypes.as<uint8_t>());
  BNX_HIP(hipGetLastError());
  BNX_SET(d.uverts, v->uverts.ptr);
  BNX_SET(d.utets, v->utets.ptr);
  BNX_SET(d.ebvh, v->ebvh.nodes.ptr);
  d.useed = nullptr;
  d.uext = nullptr;
  d.utypes = nullptr;
  if (mixed) {
    // other cells than tets: point queries go depth first through the
    // element BVH (no tet walk, no seeds)
    BNX_SET(d.uext, v->uext.ptr);
    BNX_SET(d.utypes, v->utypes.ptr);
    v->useed.release();
    // no neighbours: a walk (only in a tets-only kernel, which such a field
    // never meets) ends at its first cell
    BNX_HIP(v->unbr.ensure(sizeof(int4) * (size_t)ntet));
    BNX_HIP(hipMemsetAsync(v->unbr.ptr, 0xff, sizeof(int4) * (size_t)ntet, st));
  } else {
    // face neighbours of the tets, in slot order, for the walking point query
    if ((rc = build_tet_neighbors(dev, v->utets.as<int4>(), ntet, v->unbr)) != BNX_OK) return rc;
  }
  // the tet records: vertices and neighbours in one 80-byte record per slot
  BNX_HIP(v->utetv.ensure(sizeof(float4) * 5 * (size_t)ntet));
  k_tet_pack<<<nblk(ntet, 256), 256, 0, st>>>(v->utets.as<int4>(), v->uverts.as<float4>(), v->unbr.as<int4>(), ntet,
                                              v->utetv.as<float4>());
  BNX_HIP(hipGetLastError());
  BNX_SET(d.utetv, v->utetv.ptr);
  BNX_HIP(hipStreamSynchronize(st));
  // the records hold the neighbours; the index array stays only for the
  // other cells' vertex lists
  v->unbr.release();
  if (!mixed) {
    v->utets.release();
    d.utets = nullptr;
    // seeds for point queries (built with the element BVH in place, before
    // the seed pointer is set, so the build's own queries are depth first)
    if ((rc = build_seed_grid(v, ntet)) != BNX_OK) return rc;
  }
  // macrocell value ranges
  const int nmc = d.mc[0] * d.mc[1] * d.mc[2];
  BNX_HIP(rlo.ensure(sizeof(int) * (size_t)nmc));
  BNX_HIP(rhi.ensure(sizeof(int) * (size_t)nmc));
  BNX_HIP(v->ranges.ensure(sizeof(float2) * (size_t)nmc));
  BNX_HIP(v->majorants.ensure(sizeof(float) * (size_t)nmc));
  BNX_SET(d.ranges, v->ranges.ptr);
  BNX_SET(d.majorants, v->majorants.ptr);
  k_fill_ranges<<<nblk(nmc, 256), 256, 0, st>>>(rlo.as<int>(), rhi.as<int>(), nmc);
  k_umesh_raster<<<nblk(ntet, 128), 128, 0, st>>>(tets.as<int4>(), ext_d, typ_d, v->uverts.as<float4>(), ntet, d,
                                                  rlo.as<int>(), rhi.as<int>());
  k_finish_ranges<<<nblk(nmc, 256), 256, 0, st>>>(rlo.as<int>(), rhi.as<int>(), nmc, v->ranges.as<float2>());
  k_majorants<<<nblk(nmc, 256), 256, 0, st>>>(d, v->ranges.as<float2>(), v->majorants.as<float>());
  BNX_HIP(hipStreamSynchronize(st));
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int volume_rebuild_majorants(Volume *v)
{
  Device *dev = v->dev;
  const size_t nmc = (size_t)v->dd.mc[0] * v->dd.mc[1] * v->dd.mc[2];
  k_majorants<<<nblk((long)nmc, 256), 256, 0, dev->stream>>>(v->dd, v->ranges.as<float2>(),
                                                             v->majorants.as<float>());
  BNX_HIP(hipStreamSynchronize(dev->stream));
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

int launch_volume_march(Device *dev, const SlotDev *slots, const SegView &sv, int blocks256, const RayQueue &q,
                        bool surfaces, int full, unsigned long long *stats)
{
  if (blocks256 > 0 && (full & MARCH_CELLS))
    k_volume_march<MARCH_FULL | MARCH_CELLS><<<blocks256, 256, 0, dev->stream>>>(slots, sv, q, surfaces ? 1 : 0, stats);
  else if (blocks256 > 0 && full)
    k_volume_march<MARCH_FULL><<<blocks256, 256, 0, dev->stream>>>(slots, sv, q, surfaces ? 1 : 0, stats);
  else if (blocks256 > 0)
    k_volume_march<0><<<blocks256, 256, 0, dev->stream>>>(slots, sv, q, surfaces ? 1 : 0, stats);
  BNX_HIP(hipGetLastError());
  return BNX_OK;
}

} // namespace bnx

// bnx_volume.h -- volume and isosurface device data (structured and
// tetrahedral fields) and the per-ray march.
#pragma once

#include "bnx_internal.h"
#include "bnx_render.h"

namespace bnx {

// hit-record geometry index of a volume event: VOLUME_GEOM_BASE + volume
// slot; of an isosurface hit: ISO_GEOM_BASE + isosurface slot
constexpr int VOLUME_GEOM_BASE = 0x40000000;
constexpr int ISO_GEOM_BASE = 0x60000000;

enum { FIELD_STRUCTURED = 0, FIELD_UMESH = 1 };

// element-BVH point-query stack: a ring of PQ_SHORT entries per thread in
// LDS ([entry][thread], 256 threads per block) spilling to PQ_DEEP private
// entries; a tet mesh whose element BVH is deeper is rejected at creation
constexpr int PQ_TB = 256;
constexpr int PQ_SHORT = 16;
constexpr int PQ_DEEP = 48;

struct VolumeDev {
  int field;                  // FIELD_STRUCTURED / FIELD_UMESH
  // structured (StructuredData, volume/StructuredData.h:30-74)
  const float BNX_G *scalars;   // dims[0]*dims[1]*dims[2], x fastest
  // the same values in bricks of B^3 cells (B = 1 << brick_shift) with a
  // one-texel apron: (B+1)^3 texels per brick, x fastest inside a brick,
  // bricks x fastest; every trilinear footprint lies inside the brick of
  // its lower corner (null: sample the linear array)
  const float BNX_G *bricks;
  int brick_shift;
  int nbricks[3];
  int dims[3];
  int ncells[3];
  float origin[3];
  float rcp_spacing[3];       // 1/gridSpacing (host float division)
  // tets (UMeshField, umesh/common/UMeshField.h:52-116)
  const float4 BNX_G *uverts;   // position, scalar
  const int4 BNX_G *utets;      // vertex indices, in element-BVH slot order
  // per tet slot an 80-byte record: its four vertices (position, scalar)
  // and then its face neighbours (int4 bits: the slot across the face
  // opposite vertex k, -1 at the boundary), so that a tet test or a walk
  // step is one dependent fetch instead of an index fetch and four gathers
  const float4 BNX_G *utetv;
  // meshes with other cells than tets: per slot the cell type (CELL_*) and
  // vertex indices 4..7 (0..3 in utets); null for tets only
  const uint8_t BNX_G *utypes;
  const int4 BNX_G *uext;
  const Node BNX_G *ebvh;       // element BVH
  // seed grid over the bounds: per cell the tet slot holding its centre (-1:
  // none), where a point query with no nearby previous tet starts its walk
  const int BNX_G *useed;
  int seed_dims[3];
  float seed_origin[3];
  float seed_rcp[3];          // cells per unit length (host float division)
  // both
  float lo[3], hi[3];         // world bounds
  int mc[3];
  int cells_per_mc;
  float mc_origin[3];
  float mc_rcp_spacing[3];    // 1 / macrocell spacing (host float division)
  const float2 BNX_G *ranges;   // macrocell value ranges (lo, hi)
  const float BNX_G *majorants;
  const float4 BNX_G *tf;
  int ntf;
  float dom0, dom1, base_density;
  int rtc_inst;               // rtc instance index (seeds the march RNG)
  int user_inst;
  int user_id;
};

// IsoSurface::DD (geometry/IsoSurface.h:31-42) with its field and an
// opaque matte material
struct IsoDev {
  VolumeDev field;
  float iso;
  float color[3];
  int rtc_inst;
  int user_inst;
  int user_id;
};

struct VolumeList {
  const VolumeDev BNX_G *vols;  // device array
  int count;
};

struct Volume {
  Device *dev;
  uint64_t version = 0;         // bumped whenever dd changes
  VolumeDev dd{};
  DevMem scalars, bricks, ranges, majorants, tf;
  DevMem uverts, utets, unbr, useed, utypes, uext, utetv;
  Bvh ebvh;                     // element BVH (cells)
};

struct IsoSurface {
  Device *dev;
  Volume *field;
  float iso;
  float color[3];
  int user_id;
};

int volume_build(Volume *v);
// the structured field's bricked copy (VolumeDev::bricks) with bricks of
// 2^shift cells; shift 0 keeps the linear layout only
int volume_brick(Volume *v, int shift);
// cells: element BVH, macrocell ranges (rasterTet<5> for tets, the cell's
// box for the others), majorants.  tets4: the host copy of the cells'
// vertex indices 0..3 (num_cells x int4), ext4 of 4..7 and types the cell
// types (CELL_*; both null when every cell is a tet), verts4 of the vertices
int volume_build_umesh(Volume *v, const int4 *tets4, const int4 *ext4, const uint8_t *types, int num_cells,
                       const float4 *verts4, int num_verts);
int volume_rebuild_majorants(Volume *v);
// upload the world's volume and isosurface tables (if changed) and return them
int world_volumes(World *w, VolumeList &vl);
int world_isos(World *w, const struct IsoDev BNX_G *&isos, int &count);
// march every ray of the generation's segments (blocks256 blocks of 256
// threads) through its request's isosurfaces and volumes; surfaces: false
// when the trace kernel did not run, so the hit records are not written yet
// full: march_kind bits of the batch's volumes and isosurface fields
// the march kernel a batch needs: the lean one (structured volumes only),
// MARCH_FULL (isosurfaces or unstructured volumes) or, with MARCH_CELLS
// also set, the one whose point queries handle other cells than tets
enum { MARCH_FULL = 1, MARCH_CELLS = 2 };
inline int march_kind(const Volume *v)
{
  if (v->dd.field == FIELD_STRUCTURED) return 0;
  return MARCH_FULL | (v->dd.utypes ? MARCH_CELLS : 0);
}
int launch_volume_march(Device *dev, const SlotDev *d_slots, const SegView &sv, int blocks256, const RayQueue &q,
                        bool surfaces, int full, unsigned long long *stats);

// ---------------------------------------------------------------- device

// CPU texture LINEAR/CLAMP addressing (rtcore/cpu/Texture.cpp:324-336)
BNX_D void lerp_coord(int N, float tc, int &i0, int &i1, float &f)
{
  if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
  if (tc - .5f >= (float)(N - 1)) { i0 = i1 = N - 1; f = 0.f; return; }
  f = (tc - .5f) - (float)(int)(tc - .5f);
  i0 = (int)(tc - .5f);
  i1 = i0 + 1;
}

BNX_D float lerp_l(float f, float a, float b) { return (1.f - f) * a + f * b; }

// per-thread march state: point-query stack and work counters (samples:
// Woodcock density evaluations; cells: macrocells stepped; unodes/utets:
// element-BVH nodes and tets tested by point queries)
struct MarchCtx {
  uint32_t *ring;             // &ring[0][thread]
  uint32_t samples, cells, unodes, utets;
  // the tet that held this ray's previous point query, and its field
  const float4 BNX_G *last_field = nullptr;
  int last_tet = -1;
};

// StructuredDataSampler::DD::sample (volume/StructuredData.h:101-112)
BNX_D float structured_sample(const VolumeDev &v, v3 P)
{
  float rx = (P.x - v.origin[0]) * v.rcp_spacing[0];
  float ry = (P.y - v.origin[1]) * v.rcp_spacing[1];
  float rz = (P.z - v.origin[2]) * v.rcp_spacing[2];
  if (rx < 0.f || ry < 0.f || rz < 0.f) return __builtin_nanf("");
  if (rx >= (float)v.ncells[0] || ry >= (float)v.ncells[1] || rz >= (float)v.ncells[2]) return __builtin_nanf("");
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lerp_coord(v.dims[0], rx + .5f, x0, x1, fx);
  lerp_coord(v.dims[1], ry + .5f, y0, y1, fy);
  lerp_coord(v.dims[2], rz + .5f, z0, z1, fz);
  const float BNX_G *b00, *b01, *b10, *b11;
  if (v.bricks) {
    // the brick of the lower corner holds the whole footprint (apron)
    const int sh = v.brick_shift, E = (1 << sh) + 1;
    const int bx = min(x0 >> sh, v.nbricks[0] - 1), by = min(y0 >> sh, v.nbricks[1] - 1),
              bz = min(z0 >> sh, v.nbricks[2] - 1);
    const float BNX_G *br =
        v.bricks + (uint32_t)((bz * v.nbricks[1] + by) * v.nbricks[0] + bx) * (uint32_t)(E * E * E);
    const int lx = x0 - (bx << sh), ly0 = y0 - (by << sh), ly1 = y1 - (by << sh), lz0 = z0 - (bz << sh),
              lz1 = z1 - (bz << sh);
    b00 = br + (ly0 + E * lz0) * E - (bx << sh);
    b01 = br + (ly1 + E * lz0) * E - (bx << sh);
    b10 = br + (ly0 + E * lz1) * E - (bx << sh);
    b11 = br + (ly1 + E * lz1) * E - (bx << sh);
    (void)lx;
  } else {
    const size_t Nx = (size_t)v.dims[0], NxNy = (size_t)v.dims[0] * (size_t)v.dims[1];
    const float BNX_G *sc = v.scalars;
    b00 = sc + (size_t)y0 * Nx + (size_t)z0 * NxNy;
    b01 = sc + (size_t)y1 * Nx + (size_t)z0 * NxNy;
    b10 = sc + (size_t)y0 * Nx + (size_t)z1 * NxNy;
    b11 = sc + (size_t)y1 * Nx + (size_t)z1 * NxNy;
  }
  float v00 = lerp_l(fx, b00[x0], b00[x1]);
  float v01 = lerp_l(fx, b01[x0], b01[x1]);
  float v10 = lerp_l(fx, b10[x0], b10[x1]);
  float v11 = lerp_l(fx, b11[x0], b11[x1]);
  float v0 = lerp_l(fy, v00, v01);
  float v1 = lerp_l(fy, v10, v11);
  return lerp_l(fz, v0, v1);
}

// evalToImplicitPlane (umesh/common/UMeshField.h:365-371)
BNX_D float implicit_plane(v3 P, float4 a, float4 b, float4 c)
{
  const v3 N = cross(xyz(b) - xyz(a), xyz(c) - xyz(a));
  return dot(P - xyz(a), N);
}

// UMeshField::DD::tetScalar (umesh/common/UMeshField.h:391-436)
BNX_D bool tet_scalar(const VolumeDev &v, int slot, v3 P, float &out)
{
  const float4 BNX_G *rec = v.utetv + 5 * (size_t)slot;
  const float4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3v = rec[3];
  const float t3 = implicit_plane(P, v0, v1, v2);
  if (t3 < 0.f) return false;
  const float t2 = implicit_plane(P, v0, v3v, v1);
  if (t2 < 0.f) return false;
  const float t1 = implicit_plane(P, v0, v2, v3v);
  if (t1 < 0.f) return false;
  const float t0 = implicit_plane(P, v1, v3v, v2);
  if (t0 < 0.f) return false;
  const float scale = 1.f / (((t0 + t1) + t2) + t3);
  out = scale * (((t0 * v0.w + t1 * v1.w) + t2 * v2.w) + t3 * v3v.w);
  return true;
}

// ------------------------------------------------------ isoparametric cells
// Pyramids, prisms and hexahedra (UMeshField::DD::eltScalar and pyrScalar /
// prismScalar / hexScalar, umesh/common/UMeshField.h:372-389, 438-495, with
// the Newton inversions of ElementIntersection.h:77-476): the cell maps
// parametric (r, s, t) to x(r, s, t) = sum_i N_i X_i over its vertices X_i
// in VTK order.  P is mapped back by Newton's method from (.5, .5, .5): at
// most 10 steps, each solving J d = x - P by Cramer's rule; converged when
// every |d_k| < 1e-4; abandoned when |det J| falls below the cell's
// tolerance or a coordinate leaves [-1e6, 1e6].  P is inside when the
// coordinates lie in the parameter domain, widened by 1e-6; the value is
// sum_i N_i f_i with the N_i of the last step's starting point.
enum { CELL_TET = 0, CELL_HEX = 1, CELL_PRISM = 2, CELL_PYR = 3 };
constexpr float CELL_IN_LO = 0.f - 1e-6f, CELL_IN_HI = 1.f + 1e-6f;

template <int TYPE> struct CellShape;
// trilinear on the unit cube
template <> struct CellShape<CELL_HEX> {
  static constexpr int N = 8;
  BNX_D static void eval(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
  {
    const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
    w[0] = rm * sm * tm; w[1] = r * sm * tm; w[2] = r * s * tm; w[3] = rm * s * tm;
    w[4] = rm * sm * t;  w[5] = r * sm * t;  w[6] = r * s * t;  w[7] = rm * s * t;
    dr[0] = -sm * tm; dr[1] = sm * tm; dr[2] = s * tm; dr[3] = -s * tm;
    dr[4] = -sm * t;  dr[5] = sm * t;  dr[6] = s * t;  dr[7] = -s * t;
    ds[0] = -rm * tm; ds[1] = -r * tm; ds[2] = r * tm; ds[3] = rm * tm;
    ds[4] = -rm * t;  ds[5] = -r * t;  ds[6] = r * t;  ds[7] = rm * t;
    dt[0] = -rm * sm; dt[1] = -r * sm; dt[2] = -r * s; dt[3] = -rm * s;
    dt[4] = rm * sm;  dt[5] = r * sm;  dt[6] = r * s;  dt[7] = rm * s;
  }
  BNX_D static bool inside(float r, float s, float t)
  {
    return r >= CELL_IN_LO && r <= CELL_IN_HI && s >= CELL_IN_LO && s <= CELL_IN_HI && t >= CELL_IN_LO &&
           t <= CELL_IN_HI;
  }
};
// linear triangle (r, s) times linear t
template <> struct CellShape<CELL_PRISM> {
  static constexpr int N = 6;
  BNX_D static void eval(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
  {
    w[0] = (1.f - r - s) * (1.f - t); w[1] = r * (1.f - t); w[2] = s * (1.f - t);
    w[3] = (1.f - r - s) * t;         w[4] = r * t;         w[5] = s * t;
    dr[0] = -1.f + t; dr[1] = 1.f - t; dr[2] = 0.f; dr[3] = -t; dr[4] = t;   dr[5] = 0.f;
    ds[0] = -1.f + t; ds[1] = 0.f;     ds[2] = 1.f - t; ds[3] = -t; ds[4] = 0.f; ds[5] = t;
    dt[0] = -1.f + r + s; dt[1] = -r; dt[2] = -s; dt[3] = 1.f - r - s; dt[4] = r; dt[5] = s;
  }
  BNX_D static bool inside(float r, float s, float t)
  {
    return r >= CELL_IN_LO && r <= CELL_IN_HI && s >= CELL_IN_LO && s <= CELL_IN_HI && t >= CELL_IN_LO &&
           t <= CELL_IN_HI && r + s <= CELL_IN_HI;
  }
};
// bilinear base quad collapsing linearly in t onto the apex (vertex 4)
template <> struct CellShape<CELL_PYR> {
  static constexpr int N = 5;
  BNX_D static void eval(float r, float s, float t, float *w, float *dr, float *ds, float *dt)
  {
    const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
    w[0] = rm * sm * tm; w[1] = r * sm * tm; w[2] = r * s * tm; w[3] = rm * s * tm; w[4] = t;
    dr[0] = -(s - 1.f) * (t - 1.f); dr[1] = (s - 1.f) * (t - 1.f); dr[2] = s - s * t; dr[3] = s * (t - 1.f);
    dr[4] = 0.f;
    ds[0] = -(r - 1.f) * (t - 1.f); ds[1] = r * (t - 1.f); ds[2] = r - r * t; ds[3] = (r - 1.f) * (t - 1.f);
    ds[4] = 0.f;
    dt[0] = -(r - 1.f) * (s - 1.f); dt[1] = r * (s - 1.f); dt[2] = -r * s; dt[3] = (r - 1.f) * s; dt[4] = 1.f;
  }
  BNX_D static bool inside(float r, float s, float t) { return CellShape<CELL_HEX>::inside(r, s, t); }
};

// det [a b c] (columns) = a . (b x c)
BNX_D float det3(v3 a, v3 b, v3 c) { return dot(a, cross(b, c)); }

template <int TYPE> BNX_D bool cell_invert(const float4 *V, v3 P, float tol, float &value)
{
  using C = CellShape<TYPE>;
  float r = .5f, s = .5f, t = .5f;
  float w[C::N], dr[C::N], ds[C::N], dt[C::N];
  bool converged = false;
  for (int it = 0; it < 10 && !converged; it++) {
    C::eval(r, s, t, w, dr, ds, dt);
    v3 f{0.f, 0.f, 0.f}, cr{0.f, 0.f, 0.f}, cs{0.f, 0.f, 0.f}, ct{0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < C::N; i++) {
      const v3 X = xyz(V[i]);
      f = f + X * w[i];
      cr = cr + X * dr[i];
      cs = cs + X * ds[i];
      ct = ct + X * dt[i];
    }
    f = f - P;
    const float d = det3(cr, cs, ct);
    if (fabsf(d) < tol) return false;
    const float d0 = det3(f, cs, ct) / d;
    const float d1 = det3(cr, f, ct) / d;
    const float d2 = det3(cr, cs, f) / d;
    r = r - d0;
    s = s - d1;
    t = t - d2;
    if (fabsf(d0) < 1e-4f && fabsf(d1) < 1e-4f && fabsf(d2) < 1e-4f) converged = true;
    else if (fabsf(r) > 1e6f || fabsf(s) > 1e6f || fabsf(t) > 1e6f) return false;
  }
  if (!converged || !C::inside(r, s, t)) return false;
  float val = 0.f;
#pragma unroll
  for (int i = 0; i < C::N; i++) val += w[i] * V[i].w;
  value = val;
  return true;
}

struct CellHit {
  float value;
  int inside;
};
// a pyramid, prism or hexahedron at slot: vertex indices 0..3 in utets,
// 4..7 in uext.  Kept out of line: the tet-only march never calls it, and
// its registers then stay out of that kernel's budget.
BNX_D CellHit cell_scalar(const float4 BNX_G *uverts, const int4 BNX_G *utets,
                                                                const int4 BNX_G *uext, int type, int slot, v3 P)
{
  const int4 a = utets[slot], b = uext[slot];
  float4 V[8];
  V[0] = uverts[a.x];
  V[1] = uverts[a.y];
  V[2] = uverts[a.z];
  V[3] = uverts[a.w];
  V[4] = uverts[b.x];
  CellHit h{0.f, 0};
  bool in;
  if (type == CELL_PYR) {
    in = cell_invert<CELL_PYR>(V, P, 1e-6f, h.value);
  } else {
    V[5] = uverts[b.y];
    if (type == CELL_PRISM) {
      in = cell_invert<CELL_PRISM>(V, P, 1e-6f, h.value);
    } else {
      V[6] = uverts[b.z];
      V[7] = uverts[b.w];
      // hexahedra: the determinant tolerance scales with the cell's size
      v3 lo = xyz(V[0]), hi = lo;
      for (int i = 1; i < 8; i++) {
        lo = v3{fmn(lo.x, V[i].x), fmn(lo.y, V[i].y), fmn(lo.z, V[i].z)};
        hi = v3{fmx(hi.x, V[i].x), fmx(hi.y, V[i].y), fmx(hi.z, V[i].z)};
      }
      in = cell_invert<CELL_HEX>(V, P, 1e-10f * length(hi - lo), h.value);
    }
  }
  h.inside = in ? 1 : 0;
  return h;
}

// closed box test, branch-free (bitwise and) so that a node's loads are
// issued together rather than behind each other
BNX_D bool box_holds(float lx, float ly, float lz, float hx, float hy, float hz, v3 P)
{
  return (lx <= P.x) & (P.x <= hx) & (ly <= P.y) & (P.y <= hy) & (lz <= P.z) & (P.z <= hz);
}

// UMeshCuBQLSampler::DD::sample (umesh/mc/UMeshCuBQLSampler.h:46-62): a
// fixed-box query with the point P (closed box tests), depth first with
// child 0 before child 1, leaf slots in order; the first tet that contains
// P gives the value, NaN if none does
// The four plane values of P against the tet in 'slot' (t_k: against the face
// opposite vertex k, positive inside), their sum and the interpolated value
struct TetPlanes {
  float t[4];
  float sum;
  float value;
};
BNX_D TetPlanes tet_planes(const VolumeDev &v, int slot, v3 P)
{
  const float4 BNX_G *rec = v.utetv + 5 * (size_t)slot;
  const float4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3v = rec[3];
  TetPlanes r;
  r.t[3] = implicit_plane(P, v0, v1, v2);
  r.t[2] = implicit_plane(P, v0, v3v, v1);
  r.t[1] = implicit_plane(P, v0, v2, v3v);
  r.t[0] = implicit_plane(P, v1, v3v, v2);
  r.sum = ((r.t[0] + r.t[1]) + r.t[2]) + r.t[3];
  r.value = (1.f / r.sum) * (((r.t[0] * v0.w + r.t[1] * v1.w) + r.t[2] * v2.w) + r.t[3] * v3v.w);
  return r;
}
// P lies inside the tet by a margin: every normalized barycentric coordinate
// above 1e-4.  In a mesh of non-overlapping tets no other tet's test can
// then accept P, so the value is the one the depth-first query would
// return, bit for bit (tet_scalar's arithmetic, in the same order).
BNX_D bool planes_interior(const TetPlanes &r)
{
  const float m = 1e-4f * r.sum;
  return r.t[0] > m && r.t[1] > m && r.t[2] > m && r.t[3] > m;
}

constexpr int TET_WALK_STEPS = 6;

// Walk from tet slot `cur` toward P: test the tet, else cross the face P
// lies furthest behind, up to TET_WALK_STEPS times.  Accepts only a tet that
// holds P with margin: in a mesh of non-overlapping tets no other tet can,
// so the value is the one the depth-first query would return (see
// planes_interior).
BNX_D bool umesh_walk(const VolumeDev &v, int cur, v3 P, MarchCtx &cx, float &value, int &found)
{
  for (int step = 0; step <= TET_WALK_STEPS; step++) {
    cx.utets++;
    const TetPlanes r = tet_planes(v, cur, P);
    if (planes_interior(r)) {
      found = cur;
      value = r.value;
      return true;
    }
    if (step == TET_WALK_STEPS || !(r.sum > 0.f)) break;
    int k = 0;
    for (int j = 1; j < 4; j++)
      if (r.t[j] < r.t[k]) k = j;
    const int4 nb = ((const int4 BNX_G *)v.utetv)[5 * (size_t)cur + 4];
    const int next = k == 0 ? nb.x : (k == 1 ? nb.y : (k == 2 ? nb.z : nb.w));
    if (next < 0) break;
    cur = next;
  }
  return false;
}

// UMeshCuBQLSampler::sample (umesh/mc/UMeshCuBQLSampler.h:46-62): the first
// tet, depth first through the element BVH, that contains P; its slot (-1:
// none, the value is then NaN)
// CELLS: the mesh may hold other cells than tets (else every cell is one)
template <bool CELLS = false> BNX_D int umesh_dfs(const VolumeDev &v, v3 P, MarchCtx &cx, float &value)
{
  uint32_t deep[PQ_DEEP];
  int sp = 0;
  uint32_t cur = 0;
  while (true) {
    if (!(cur & LEAF_BIT)) {
      const float4 BNX_G *np = (const float4 BNX_G *)(v.ebvh + cur);
      const float4 a = np[0], b = np[1], c = np[2];
      const int4 r = ((const int4 BNX_G *)np)[3];
      cx.unodes++;
      const bool h0 = (r.z >= 0) & box_holds(a.x, a.y, a.z, a.w, b.x, b.y, P);
      const bool h1 = (r.w >= 0) & box_holds(b.z, b.w, c.x, c.y, c.z, c.w, P);
      if (h0 && h1) {
        // push child 1
        const int s = sp & (PQ_SHORT - 1);
        if (sp >= PQ_SHORT) deep[sp - PQ_SHORT] = cx.ring[s * PQ_TB];
        cx.ring[s * PQ_TB] = enc_child(r.y, r.w);
        sp++;
        cur = enc_child(r.x, r.z);
        continue;
      }
      if (h0) { cur = enc_child(r.x, r.z); continue; }
      if (h1) { cur = enc_child(r.y, r.w); continue; }
    } else {
      const int first = (int)((cur & ~LEAF_BIT) >> 3);
      const int count = (int)(cur & 7u);
      for (int k = 0; k < count; k++) {
        float f;
        cx.utets++;
        bool in;
        const int type = CELLS && v.utypes ? (int)v.utypes[first + k] : CELL_TET;
        if (type == CELL_TET) {
          in = tet_scalar(v, first + k, P, f);
        } else {
          const CellHit h = cell_scalar(v.uverts, v.utets, v.uext, type, first + k, P);
          in = h.inside != 0;
          f = h.value;
        }
        if (in) {
          value = f;
          return first + k;
        }
      }
    }
    if (sp == 0) {
      value = __builtin_nanf("");
      return -1;
    }
    sp--;
    const int s = sp & (PQ_SHORT - 1);
    cur = cx.ring[s * PQ_TB];
    if (sp >= PQ_SHORT) cx.ring[s * PQ_TB] = deep[sp - PQ_SHORT];
  }
}

// the seed-grid cell of P (clamped to the grid)
BNX_D int seed_cell(const VolumeDev &v, v3 P)
{
  const float p3[3] = {P.x, P.y, P.z};
  int c[3];
  for (int k = 0; k < 3; k++) {
    const float x = fminf(fmaxf((p3[k] - v.seed_origin[k]) * v.seed_rcp[k], 0.f), (float)(v.seed_dims[k] - 1));
    c[k] = (int)x;
  }
  return (c[2] * v.seed_dims[1] + c[1]) * v.seed_dims[0] + c[0];
}

// A ray's consecutive queries are close together: walk from the tet of its
// previous query; a query with none nearby walks from the seed of its grid
// cell; the depth-first query decides what neither walk settles.
template <bool CELLS = false> BNX_D float umesh_sample(const VolumeDev &v, v3 P, MarchCtx &cx)
{
  float value;
  int t;
  if (CELLS && v.utypes) {
    // other cells than tets: the depth-first query alone, no walk (a kernel
    // compiled for tets only is never launched with such a field)
    (void)umesh_dfs<CELLS>(v, P, cx, value);
    return value;
  }
  const bool have_last = cx.last_field == v.utetv && cx.last_tet >= 0;
  if (have_last && umesh_walk(v, cx.last_tet, P, cx, value, t)) {
    cx.last_tet = t;
    return value;
  }
  if (v.useed) {
    const int seed = v.useed[seed_cell(v, P)];
    if (seed >= 0 && umesh_walk(v, seed, P, cx, value, t)) {
      cx.last_field = v.utetv;
      cx.last_tet = t;
      return value;
    }
  }
  t = umesh_dfs(v, P, cx, value);
  if (t >= 0) {
    cx.last_field = v.utetv;
    cx.last_tet = t;
  }
  return value;
}

template <bool CELLS = false> BNX_D float volume_sample(const VolumeDev &v, v3 P, MarchCtx &cx)
{
  return v.field == FIELD_UMESH ? umesh_sample<CELLS>(v, P, cx) : structured_sample(v, P);
}

// TransferFunction::DD::map (volume/TransferFunction.h:67-87)
BNX_D float4 tf_map(const VolumeDev &v, float s)
{
  const int N = v.ntf;
  float f = (s - v.dom0) / (v.dom1 - v.dom0);
  f = fmn(fmx(f, 0.f), 1.f);
  f *= (float)(N - 1);
  int idx = (int)f;
  idx = idx < 0 ? 0 : (idx > N - 2 ? N - 2 : idx);
  f -= (float)idx;
  const float4 BNX_G *tf = v.tf;
  float4 a = tf[idx], b = tf[idx + 1];
  float4 r;
  r.x = (1.f - f) * a.x + f * b.x;
  r.y = (1.f - f) * a.y + f * b.y;
  r.z = (1.f - f) * a.z + f * b.z;
  r.w = (1.f - f) * a.w + f * b.w;
  r.w *= v.base_density;
  return r;
}

// TransferFunction::DD::majorant (volume/TransferFunction.h:89-117)
BNX_D float tf_majorant(const VolumeDev &v, float lo, float hi)
{
  if (lo > hi) return 0.f;
  const int N = v.ntf;
  const float span = v.dom1 - v.dom0;
  float f_lo = (lo - v.dom0) / span;
  float f_hi = (hi - v.dom0) / span;
  f_lo = fmn(fmx(f_lo, 0.f), 1.f);
  f_hi = fmn(fmx(f_hi, 0.f), 1.f);
  f_lo *= (float)(N - 1);
  f_hi *= (float)(N - 1);
  int i_lo = (int)f_lo, i_hi = (int)f_hi;
  i_lo = i_lo < 0 ? 0 : (i_lo > N - 2 ? N - 2 : i_lo);
  i_hi = i_hi < 0 ? 0 : (i_hi > N - 2 ? N - 2 : i_hi);
  f_lo -= (float)i_lo;
  f_hi -= (float)i_hi;
  const float4 BNX_G *tf = v.tf;
  float l_lo = (1.f - f_lo) * tf[i_lo].w + f_lo * tf[i_lo + 1].w;
  float l_hi = (1.f - f_hi) * tf[i_hi].w + f_hi * tf[i_hi + 1].w;
  float m = 0.f;
  m = fmx(m, l_lo);
  m = fmx(m, l_hi);
  for (int i = i_lo + 1; i <= i_hi; i++) m = fmx(m, tf[i].w);
  return m * v.base_density;
}

// fastLog (volume/Volume.h:154-165)
BNX_D float fast_log(float f)
{
  f = (f - 1.f) / (f + 1.f);
  float f2 = f * f;
  float s = f;
  f *= f2;
  s += (1.f / 3.f) * f;
  f *= f2;
  s += (1.f / 5.f) * f;
  return s + s;
}

// boxTest (render/Ray.h:169-181): clip [r0, r1] to the field's bounds
BNX_D bool clip_bounds(const VolumeDev &v, v3 org, v3 dir, float &r0, float &r1)
{
  v3 rd = v3{1.f / dir.x, 1.f / dir.y, 1.f / dir.z};
  float tl[3] = {(v.lo[0] - org.x) * rd.x, (v.lo[1] - org.y) * rd.y, (v.lo[2] - org.z) * rd.z};
  float th[3] = {(v.hi[0] - org.x) * rd.x, (v.hi[1] - org.y) * rd.y, (v.hi[2] - org.z) * rd.z};
  float nr = fmx(fmx(fmn(tl[0], th[0]), fmn(tl[1], th[1])), fmn(tl[2], th[2]));
  float fr = fmn(fmn(fmx(tl[0], th[0]), fmx(tl[1], th[1])), fmx(tl[2], th[2]));
  r0 = fmx(r0, nr);
  r1 = fmn(r1, fr);
  return r0 <= r1;
}

// dda3 (volume/DDA.h:60-201): walks the macrocells of v along the ray
// (macrocell-grid space) from max(0, grid entry) to min(dda_tmax, grid
// exit), calling fn(cell index, t0, t1); stops when fn returns false
template <typename F> BNX_D void dda3(const VolumeDev &v, v3 org, v3 dir, float dda_tmax, F &&fn)
{
  const v3 dorg = v3{(org.x - v.mc_origin[0]) * v.mc_rcp_spacing[0], (org.y - v.mc_origin[1]) * v.mc_rcp_spacing[1],
                     (org.z - v.mc_origin[2]) * v.mc_rcp_spacing[2]};
  const v3 ddir = v3{dir.x * v.mc_rcp_spacing[0], dir.y * v.mc_rcp_spacing[1], dir.z * v.mc_rcp_spacing[2]};
  const float fs[3] = {(float)v.mc[0], (float)v.mc[1], (float)v.mc[2]};
  const float o3[3] = {dorg.x, dorg.y, dorg.z}, d3[3] = {ddir.x, ddir.y, ddir.z};
  float rcp[3], tn[3], tf[3];
  for (int k = 0; k < 3; k++) {
    rcp[k] = 1.f / d3[k];
    float a = (0.f - o3[k]) * rcp[k];
    float b = (fs[k] - o3[k]) * rcp[k];
    tn[k] = fmn(a, b);
    tf[k] = fmx(a, b);
    if (d3[k] == 0.f) {
      if (o3[k] < 0.f || o3[k] > fs[k]) return;
      tn[k] = -INFINITY;
      tf[k] = INFINITY;
    }
  }
  float ray_t0 = fmx(0.f, fmx(fmx(tn[0], tn[1]), tn[2]));
  float ray_t1 = fmn(dda_tmax, fmn(fmn(tf[0], tf[1]), tf[2]));
  if (ray_t0 > ray_t1) return;
  float t_step[3], t_next[3];
  int stop[3], delta[3], cell[3];
  for (int k = 0; k < 3; k++) {
    float oiv = o3[k] + ray_t0 * d3[k];
    float fc = fmx(0.f, fmn(fs[k] - 1.f, floorf(oiv)));
    float fe = d3[k] > 0.f ? fc + 1.f : fc;
    t_step[k] = fabsf(rcp[k]);
    t_next[k] = (d3[k] == 0.f) ? INFINITY : fabsf(fe - oiv) * t_step[k];
    stop[k] = d3[k] > 0.f ? v.mc[k] : -1;
    delta[k] = d3[k] > 0.f ? 1 : -1;
    cell[k] = (int)fc;
  }
  float next_begin = 0.f;
  while (true) {
    const float t_closest = fmn(fmn(t_next[0], t_next[1]), t_next[2]);
    const float c0 = ray_t0 + next_begin;
    const float c1 = ray_t0 + fmn(t_closest, dda_tmax);
    const size_t ci = (size_t)cell[0] + (size_t)v.mc[0] * ((size_t)cell[1] + (size_t)v.mc[1] * cell[2]);
    if (!fn(ci, c0, c1)) return;
    next_begin = t_closest;
    if (t_next[0] == t_closest) {
      t_next[0] += t_step[0];
      cell[0] += delta[0];
      if (cell[0] == stop[0]) return;
    }
    if (t_next[1] == t_closest) {
      t_next[1] += t_step[1];
      cell[1] += delta[1];
      if (cell[1] == stop[1]) return;
    }
    if (t_next[2] == t_closest) {
      t_next[2] += t_step[2];
      cell[2] += delta[2];
      if (cell[2] == stop[2]) return;
    }
  }
}

// the ray's RNG child for one intersection program: Random(ray.rngSeed,
// hash(inst, geom, prim)) -- barney-common.h:59-61; advances the parent
BNX_D uint64_t rng_child(uint64_t &parent, uint32_t rtc_inst)
{
  uint64_t rs = parent;
  rng_next(parent);
  rng_next_h(rs, fnv3(rtc_inst, 0u, 0u));
  return rs;
}

// MCVolumeAccel::isProg (volume/MCAccelerator.h:550-689): box clip, RNG
// child seeded by hash(rtc instance, geometry 0, 0), dda3 over macrocells,
// Woodcock sampling per non-empty cell.  'dir' is the fixed trace direction.
// ANY_FIELD: the field may be unstructured (else it is structured); CELLS:
// an unstructured field may hold other cells than tets.
template <bool ANY_FIELD, bool CELLS = false>
BNX_D bool volume_march(const VolumeDev &v, v3 org, v3 dir, float tmin, float tmax, uint64_t &parent, float &t_out,
                        float rgb[3], MarchCtx &cx)
{
  if (!ANY_FIELD && v.field != FIELD_STRUCTURED) return false;   // never launched that way
  float r0 = tmin, r1 = tmax;
  if (!clip_bounds(v, org, dir, r0, r1)) return false;
  uint64_t rs = rng_child(parent, (uint32_t)v.rtc_inst);
  auto rnd = [&]() -> float { return (float)(rng_next(rs) & 0x00FFFFFFu) * (1.f / (float)0x01000000); };
  const float ray_tmax = tmax;
  bool found = false;
  dda3(v, org, dir, r1, [&](size_t ci, float c0, float c1) -> bool {
    const float maj = v.majorants[ci];
    cx.cells++;
    if (maj == 0.f) return true;
    // Woodcock::sampleRange (volume/Volume.h:168-203)
    const float upper = fmn(c1, ray_tmax);
    float t = c0;
    while (true) {
      float r = rnd();
      float dt = -fast_log(1.f - r) / maj;
      t += dt;
      if (t >= upper) return true;
      v3 P = org + t * dir;
      float f = ANY_FIELD ? volume_sample<CELLS>(v, P, cx) : structured_sample(v, P);
      cx.samples++;
      float4 s = isnan(f) ? make_float4(0.f, 0.f, 0.f, 0.f) : tf_map(v, f);
      if (s.w >= rnd() * maj) {
        t_out = t;
        rgb[0] = s.x;
        rgb[1] = s.y;
        rgb[2] = s.z;
        found = true;
        return false;
      }
    }
  });
  return found;
}

// Ratio tracking (Novak, Selle and Jarosz, "Residual Ratio Tracking for
// Estimating Attenuation in Participating Media", SIGGRAPH Asia 2014,
// section 3) of a shadow ray through one volume: the same macrocell walk,
// RNG child and tentative collisions as volume_march (one draw per step
// instead of two), and at each tentative collision the estimate is
// multiplied by the null-collision probability 1 - sigma / majorant instead
// of ending the ray with that probability.  Returns the transmittance
// estimate in [0, 1]; its expectation is the volume's transmittance along
// [tmin, tmax], as the expected value of delta tracking's 0/1 outcome is.
// The reference has only delta tracking (volume/Volume.h:168-203): this is
// an opt-in estimator (BNX_RENDER_RATIO_TRACKING).
template <bool ANY_FIELD, bool CELLS = false>
BNX_D float ratio_track(const VolumeDev &v, v3 org, v3 dir, float tmin, float tmax, uint64_t &parent, MarchCtx &cx)
{
  if (!ANY_FIELD && v.field != FIELD_STRUCTURED) return 1.f;
  float r0 = tmin, r1 = tmax;
  if (!clip_bounds(v, org, dir, r0, r1)) return 1.f;
  uint64_t rs = rng_child(parent, (uint32_t)v.rtc_inst);
  auto rnd = [&]() -> float { return (float)(rng_next(rs) & 0x00FFFFFFu) * (1.f / (float)0x01000000); };
  const float ray_tmax = tmax;
  float T = 1.f;
  dda3(v, org, dir, r1, [&](size_t ci, float c0, float c1) -> bool {
    const float maj = v.majorants[ci];
    cx.cells++;
    if (maj == 0.f) return true;
    const float upper = fmn(c1, ray_tmax);
    float t = c0;
    while (true) {
      const float r = rnd();
      t += -fast_log(1.f - r) / maj;
      if (t >= upper) return true;
      const v3 P = org + t * dir;
      const float f = ANY_FIELD ? volume_sample<CELLS>(v, P, cx) : structured_sample(v, P);
      cx.samples++;
      const float density = isnan(f) ? 0.f : tf_map(v, f).w;
      T *= fmx(0.f, 1.f - density / maj);
      if (T == 0.f) return false;
    }
  });
  return T;
}

// MCIsoSurfaceAccel::isProg (volume/MCAccelerator.h:325-548): DDA over the
// field's macrocell value ranges; in every cell whose range holds the iso
// value, 10 linear sub-steps with a sample at each end; the first crossing
// found before tmax ends the walk.  Normal: central differences with delta
// = 0.1 |bounds| / (sum of grid dims).  Advances the parent RNG once when
// the ray overlaps the field's bounds.  A crossing whose interpolation is
// NaN (both ends exactly at the iso value) is ignored.
template <bool CELLS = false>
BNX_D bool iso_march(const IsoDev &s, v3 org, v3 dir, float tmin, float tmax, uint64_t &parent, float &t_out,
                     v3 &normal, MarchCtx &cx)
{
  const VolumeDev &v = s.field;
  float r0 = tmin, r1 = tmax;
  if (!clip_bounds(v, org, dir, r0, r1)) return false;
  (void)rng_child(parent, (uint32_t)s.rtc_inst);   // drawn from only for opacity < 1
  const float ray_tmax = tmax;
  const float iso = s.iso;
  float t_hit = ray_tmax;
  dda3(v, org, dir, r1, [&](size_t ci, float t0, float t1) -> bool {
    cx.cells++;
    if (!(t0 < fmn(t1, ray_tmax))) return true;
    const float2 vr = v.ranges[ci];
    if (!(iso >= vr.x && iso <= vr.y)) return true;
    float tt1 = t0;
    float ff1 = volume_sample<CELLS>(v, org + tt1 * dir, cx);
    for (int i = 1; i <= 10; i++) {
      const float tt0 = tt1;
      const float ff0 = ff1;
      const float fi = (float)i / 10.f;
      tt1 = (1.f - fi) * t0 + fi * t1;
      ff1 = volume_sample<CELLS>(v, org + tt1 * dir, cx);
      if (isnan(ff0) || isnan(ff1)) continue;
      const float lo = fmn(ff0, ff1), hi = fmx(ff0, ff1);
      if (iso >= lo && iso <= hi) {
        float t = (iso - ff0) / (ff1 - ff0);
        t = (1.f - t) * tt0 + t * tt1;
        if (t < t_hit) t_hit = t;
        if (t_hit < ray_tmax) return false;
      }
    }
    return true;
  });
  if (!(t_hit < ray_tmax)) return false;
  const v3 P = org + t_hit * dir;
  const v3 bs = v3{v.hi[0] - v.lo[0], v.hi[1] - v.lo[1], v.hi[2] - v.lo[2]};
  const float delta = length(bs) * .1f / (float)(v.mc[0] + v.mc[1] + v.mc[2]);
  const float fP = volume_sample<CELLS>(v, P, cx);
  float fx0 = volume_sample<CELLS>(v, P + v3{-delta, 0.f, 0.f}, cx);
  float fx1 = volume_sample<CELLS>(v, P + v3{+delta, 0.f, 0.f}, cx);
  float fy0 = volume_sample<CELLS>(v, P + v3{0.f, -delta, 0.f}, cx);
  float fy1 = volume_sample<CELLS>(v, P + v3{0.f, +delta, 0.f}, cx);
  float fz0 = volume_sample<CELLS>(v, P + v3{0.f, 0.f, -delta}, cx);
  float fz1 = volume_sample<CELLS>(v, P + v3{0.f, 0.f, +delta}, cx);
  float dx = 2.f, dy = 2.f, dz = 2.f;
  if (isnan(fx0)) { dx -= 1.f; fx0 = fP; }
  if (isnan(fx1)) { dx -= 1.f; fx1 = fP; }
  if (isnan(fy0)) { dy -= 1.f; fy0 = fP; }
  if (isnan(fy1)) { dy -= 1.f; fy1 = fP; }
  if (isnan(fz0)) { dz -= 1.f; fz0 = fP; }
  if (isnan(fz1)) { dz -= 1.f; fz1 = fP; }
  v3 n = v3{dx == 0.f ? 0.f : (fx1 - fx0) / dx, dy == 0.f ? 0.f : (fy1 - fy0) / dy,
            dz == 0.f ? 0.f : (fz1 - fz0) / dz};
  if (n == splat(0.f)) n = -normalize(dir);
  t_out = t_hit;
  normal = normalize(n);
  return true;
}

} // namespace bnx

// dda.h -- CPU oracle: the macrocell-grid walk shared by the volume and the
// isosurface intersection programs.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#pragma once

#include "oracle.h"

namespace orc {

// dda3 (volume/DDA.h:60-201): walks the cells of a dims[0] x dims[1] x
// dims[2] grid along dorg + t * ddir (grid space) from max(0, grid entry)
// to min(dda_tmax, grid exit), calling fn(cell, t0, t1) with the cell's
// parametric range (t0 = entry + previous crossing, t1 = entry + min(next
// crossing, dda_tmax) -- the reference's bookkeeping); stops when fn returns
// false or the walk leaves the grid.
template <typename F> void dda3(vec3 dorg, vec3 ddir, float dda_tmax, const int dims[3], F &&fn)
{
  const vec3 f_size((float)dims[0], (float)dims[1], (float)dims[2]);
  const vec3 rcp_dir(1.f / ddir.x, 1.f / ddir.y, 1.f / ddir.z);
  vec3 t_lo = (vec3(0.f) - dorg) * rcp_dir;
  vec3 t_hi = (f_size - dorg) * rcp_dir;
  vec3 t_nr = vmin(t_lo, t_hi), t_fr = vmax(t_lo, t_hi);
  for (int k = 0; k < 3; k++) {
    if (ddir[k] == 0.f) {
      if (dorg[k] < 0.f || dorg[k] > f_size[k]) return;
      t_nr[k] = -INFINITY;
      t_fr[k] = INFINITY;
    }
  }
  float ray_t0 = fmx(0.f, reduce_max(t_nr));
  float ray_t1 = fmn(dda_tmax, reduce_min(t_fr));
  if (ray_t0 > ray_t1) return;
  vec3 oiv = dorg + ray_t0 * ddir;
  vec3 f_cell;
  for (int k = 0; k < 3; k++) f_cell[k] = fmx(0.f, fmn(f_size[k] - 1.f, floorf(oiv[k])));
  vec3 f_end, t_step, t_next;
  int stop[3], delta[3], cell[3];
  for (int k = 0; k < 3; k++) {
    f_end[k] = ddir[k] > 0.f ? f_cell[k] + 1.f : f_cell[k];
    t_step[k] = fabsf(rcp_dir[k]);
    t_next[k] = (ddir[k] == 0.f) ? INFINITY : fabsf(f_end[k] - oiv[k]) * t_step[k];
    stop[k] = ddir[k] > 0.f ? dims[k] : -1;
    delta[k] = ddir[k] > 0.f ? 1 : -1;
    cell[k] = (int)f_cell[k];
  }
  float next_begin = 0.f;
  while (true) {
    float t_closest = reduce_min(t_next);
    const float c0 = ray_t0 + next_begin;
    const float c1 = ray_t0 + fmn(t_closest, dda_tmax);
    if (!fn(cell, c0, c1)) return;
    next_begin = t_closest;
    for (int k = 0; k < 3; k++) {
      if (t_next[k] == t_closest) {
        t_next[k] += t_step[k];
        cell[k] += delta[k];
        if (cell[k] == stop[k]) return;
      }
    }
  }
}

inline size_t mc_index(const Volume &v, const int cell[3])
{
  return (size_t)cell[0] + (size_t)v.mc[0] * ((size_t)cell[1] + (size_t)v.mc[1] * cell[2]);
}

} // namespace orc

// Lattice SDF query and marching cubes (gfx950): the mesh extraction of
// Voxurf.extract_geometry (lib/voxurf_fine.py:894-910 -> lib/dvgo_ori.py:679-703), whose
// reference runs the SDF query chunk by chunk through grid_sample, copies the field to
// the host and calls PyMCubes on the CPU.
//
// Here the field stays on the device: vx_extract_fields evaluates the trilinear SDF at
// the lattice points (torch.linspace's arithmetic), and marching cubes runs as two
// passes over lattice points with a scan between them: count (crossing edges owned by
// each point, triangles of the cube at each point), then emit (vertices on the owned
// edges, triangles by table).  Every vertex sits on exactly one lattice edge and is
// shared by all triangles that use that edge, so the mesh is indexed without a hash.
//
// The triangle table is generated here (vx_mc_tables), not transcribed: for each of the
// 256 corner configurations the isosurface crosses each cube face in one segment, or in
// two on a face whose below-threshold corners are diagonal (these are separated, a rule
// that depends only on the face, so neighbouring cubes agree and the mesh is closed);
// segments are oriented so that triangles face the below-threshold side, linked into
// loops around the cube, and each loop is triangulated as a fan.
#include "voxel.h"

#include <string.h>

namespace vx {

// Corner c of a cube: offsets (c & 1, c >> 1 & 1, c >> 2 & 1) along (x, y, z).
// Edge e = 4 * axis + (b0 | b1 << 1): along `axis`, at offsets b0, b1 on the other two
// axes in increasing axis order; it starts at the corner with offset 0 along `axis`.
constexpr int kMcMaxTris = 8;

struct McTables {
  int8_t count[256];
  int8_t edges[256][3 * kMcMaxTris];
};

static void edge_corners(int e, int& c0, int& c1) {
  const int axis = e >> 2, b0 = e & 1, b1 = e >> 1 & 1;
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;  // the other two axes
  c0 = (b0 << o0) | (b1 << o1);
  c1 = c0 | (1 << axis);
}

static int edge_between(int ca, int cb) {
  const int d = ca ^ cb;
  const int axis = d == 1 ? 0 : (d == 2 ? 1 : 2);
  const int base = ca & ~d;
  const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  return 4 * axis + ((base >> o0 & 1) | (base >> o1 & 1) << 1);
}

static void edge_mid(int e, double p[3]) {
  int c0, c1;
  edge_corners(e, c0, c1);
  for (int a = 0; a < 3; ++a) p[a] = 0.5 * ((c0 >> a & 1) + (c1 >> a & 1));
}

// Bit (2 * axis + side) is set for each cube face the edge lies on.
static int edge_faces(int e) {
  const int axis = e >> 2, o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
  return 1 << (2 * o0 + (e & 1)) | 1 << (2 * o1 + (e >> 1 & 1));
}

static bool same_face(int a, int b, int c) { return (edge_faces(a) & edge_faces(b) & edge_faces(c)) != 0; }

static void build_tables(McTables& t) {
  memset(&t, 0, sizeof(t));
  for (int cfg = 0; cfg < 256; ++cfg) {
    int seg_from[12], seg_to_of[12];  // segment starting at edge e ends at seg_to_of[e]
    for (int e = 0; e < 12; ++e) seg_from[e] = seg_to_of[e] = -1;
    for (int axis = 0; axis < 3; ++axis) {
      const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
      for (int side = 0; side < 2; ++side) {
        // face corners in cyclic order (u, v) = (0,0), (1,0), (1,1), (0,1)
        int fc[4];
        const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int q = 0; q < 4; ++q) fc[q] = (side << axis) | (uv[q][0] << u) | (uv[q][1] << v);
        bool below[4];
        for (int q = 0; q < 4; ++q) below[q] = cfg >> fc[q] & 1;
        int cross[4], nc = 0;
        for (int q = 0; q < 4; ++q)
          if (below[q] != below[(q + 1) & 3]) cross[nc++] = q;  // edge fc[q] - fc[q+1]
        if (nc == 0) continue;
        // segments as pairs of face-edge slots, with the below corner each cuts off
        int segs[2][2], cut[2], ns = 0;
        if (nc == 2) {
          segs[0][0] = cross[0];
          segs[0][1] = cross[1];
          int qb = -1;
          for (int q = 0; q < 4; ++q)
            if (below[q]) qb = q;
          cut[0] = qb;
          ns = 1;
        } else {
          // diagonal below corners: each is cut off by the segment of its two edges
          for (int q = 0; q < 4; ++q) {
            if (!below[q]) continue;
            segs[ns][0] = (q + 3) & 3;  // edge fc[q-1] - fc[q]
            segs[ns][1] = q;            // edge fc[q] - fc[q+1]
            cut[ns] = q;
            ++ns;
          }
        }
        double n[3] = {0, 0, 0};
        n[axis] = side ? 1.0 : -1.0;  // outward face normal
        for (int s = 0; s < ns; ++s) {
          int ea = edge_between(fc[segs[s][0]], fc[(segs[s][0] + 1) & 3]);
          int eb = edge_between(fc[segs[s][1]], fc[(segs[s][1] + 1) & 3]);
          double pa[3], pb[3], q[3];
          edge_mid(ea, pa);
          edge_mid(eb, pb);
          for (int a = 0; a < 3; ++a) q[a] = fc[cut[s]] >> a & 1;
          const double d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
          const double w[3] = {d[1] * n[2] - d[2] * n[1], d[2] * n[0] - d[0] * n[2],
                               d[0] * n[1] - d[1] * n[0]};
          const double side_q = w[0] * (q[0] - pa[0]) + w[1] * (q[1] - pa[1]) + w[2] * (q[2] - pa[2]);
          if (side_q > 0) {  // (b - a) x n must point away from the below corner
            const int tmp = ea;
            ea = eb;
            eb = tmp;
          }
          seg_from[ea] = ea;
          seg_to_of[ea] = eb;
        }
      }
    }
    // link segments into loops and fan-triangulate each
    bool used[12] = {false};
    int nt = 0;
    for (int e0 = 0; e0 < 12; ++e0) {
      if (seg_from[e0] < 0 || used[e0]) continue;
      int loop[12], nl = 0, e = e0;
      do {
        used[e] = true;
        loop[nl++] = e;
        e = seg_to_of[e];
      } while (e != e0 && e >= 0 && nl < 12);
      // fan from the first apex whose triangles all cut through the cube (no triangle
      // with its three vertices on one cube face, which a neighbouring cube could
      // duplicate in the shared face)
      int apex = 0;
      for (int a = 0; a < nl; ++a) {
        bool ok = true;
        for (int k = 1; k + 1 < nl && ok; ++k)
          ok = !same_face(loop[a], loop[(a + k) % nl], loop[(a + k + 1) % nl]);
        if (ok) {
          apex = a;
          break;
        }
      }
      for (int k = 1; k + 1 < nl; ++k) {
        t.edges[cfg][3 * nt + 0] = (int8_t)loop[apex];
        t.edges[cfg][3 * nt + 1] = (int8_t)loop[(apex + k) % nl];
        t.edges[cfg][3 * nt + 2] = (int8_t)loop[(apex + k + 1) % nl];
        ++nt;
      }
    }
    t.count[cfg] = (int8_t)nt;
  }
}

static const McTables& tables() {
  static const McTables t = [] {
    McTables x;
    build_tables(x);
    return x;
  }();
  return t;
}

// ---- lattice SDF query (lib/dvgo_ori.py:679-693 with Voxurf.grid_sampler) ------------

// torch.linspace(start, end, n)[i] in float: start + step * i for the first half,
// end - step * (n - 1 - i) for the rest, step = (end - start) / (n - 1), each value
// with one rounding (a fused multiply-add, as torch's vectorised kernel computes it).
__device__ __forceinline__ float linspace_at(float start, float end, int n, int i) {
#pragma clang fp contract(off)
  if (n == 1) return start;
  const float step = (end - start) / (float)(n - 1);
  return i < n / 2 ? fmaf(step, (float)i, start) : fmaf(-step, (float)(n - 1 - i), end);
}

struct FieldArgs {
  const float* sdf;
  GridDesc g;
  float mn[3], mx[3];   // grid box
  float bmin[3], bmax[3];
  int n[3];             // lattice points per axis
  float scale;          // -1: the field of -sdf (Voxurf.extract_geometry)
  float* u;             // [n0, n1, n2]
};

// One thread per lattice point; 3-D grid (z chunks of 256, y, x).
__global__ void __launch_bounds__(256) k_extract_fields(FieldArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
  if (k >= a.n[2]) return;
  const float px = linspace_at(a.bmin[0], a.bmax[0], a.n[0], i);
  const float py = linspace_at(a.bmin[1], a.bmax[1], a.n[1], j);
  const float pz = linspace_at(a.bmin[2], a.bmax[2], a.n[2], k);
  // the trilinear lookup of the drop-in grid_sample (voxel.h), zero outside the grid
  Tri<float> t;
  t.init(world_to_index(px, a.mn[0], a.mx[0] - a.mn[0], a.g.X),
         world_to_index(py, a.mn[1], a.mx[1] - a.mn[1], a.g.Y),
         world_to_index(pz, a.mn[2], a.mx[2] - a.mn[2], a.g.Z));
  const float v = tri_gather(a.sdf, a.g, t, 0);
  a.u[((int64_t)i * a.n[1] + j) * a.n[2] + k] = a.scale * v;
}

// ---- marching cubes ---------------------------------------------------------------------

struct McArgs {
  const float* u;
  int nx, ny, nz;
  float iso;
  const int8_t* count;  // [256] device copy of the table
  const int8_t* edges;  // [256][3 * kMcMaxTris]
};

__device__ __forceinline__ bool below(const McArgs& a, int64_t p) { return a.u[p] < a.iso; }

// Crossing bits (x, y, z) of the three lattice edges starting at point (i, j, k).
__device__ __forceinline__ int owned_bits(const McArgs& a, int i, int j, int k, int64_t p) {
  const int64_t sx = (int64_t)a.ny * a.nz, sy = a.nz;
  const bool b = below(a, p);
  int bits = 0;
  if (i + 1 < a.nx && below(a, p + sx) != b) bits |= 1;
  if (j + 1 < a.ny && below(a, p + sy) != b) bits |= 2;
  if (k + 1 < a.nz && below(a, p + 1) != b) bits |= 4;
  return bits;
}

__device__ __forceinline__ int cube_index(const McArgs& a, int64_t p) {
  const int64_t sx = (int64_t)a.ny * a.nz, sy = a.nz;
  int idx = 0;
  for (int c = 0; c < 8; ++c)
    if (below(a, p + (c & 1) * sx + (c >> 1 & 1) * sy + (c >> 2 & 1))) idx |= 1 << c;
  return idx;
}

// Pass 1: vertices owned and triangles of the cube at each lattice point.
__global__ void __launch_bounds__(256) k_mc_count(McArgs a, int64_t* __restrict__ vcnt,
                                                  int64_t* __restrict__ tcnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
  if (k >= a.nz) return;
  const int64_t p = ((int64_t)i * a.ny + j) * a.nz + k;
  vcnt[p] = __popc(owned_bits(a, i, j, k, p));
  tcnt[p] = (i + 1 < a.nx && j + 1 < a.ny && k + 1 < a.nz) ? a.count[cube_index(a, p)] : 0;
}

// Pass 2: vertices (index space, or world space when bmin/bmax are given) and faces.
__global__ void __launch_bounds__(256) k_mc_emit(McArgs a, const int64_t* __restrict__ vincl,
                                                 const int64_t* __restrict__ tincl, int world,
                                                 double b0, double b1, double b2, double e0,
                                                 double e1, double e2, double* __restrict__ verts,
                                                 int64_t* __restrict__ faces) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
  if (k >= a.nz) return;
  const int64_t sx = (int64_t)a.ny * a.nz, sy = a.nz;
  const int64_t p = (i * sx + j * sy) + k;
  const int bits = owned_bits(a, i, j, k, p);
  if (bits) {
    int64_t vid = p ? vincl[p - 1] : 0;
    const double f1 = a.u[p];
    const int64_t step[3] = {sx, sy, 1};
    const double bmin[3] = {b0, b1, b2}, ext[3] = {e0, e1, e2};
    const int nn[3] = {a.nx, a.ny, a.nz};
    for (int ax = 0; ax < 3; ++ax) {
      if (!(bits >> ax & 1)) continue;
      const double f2 = a.u[p + step[ax]];
      double v[3] = {(double)i, (double)j, (double)k};
      v[ax] += ((double)a.iso - f1) / (f2 - f1);
      for (int c = 0; c < 3; ++c) {
        double w = v[c];
        if (world) w = w / (double)(nn[c] - 1) * ext[c] + bmin[c];
        verts[3 * vid + c] = w;
      }
      ++vid;
    }
  }
  if (i + 1 < a.nx && j + 1 < a.ny && k + 1 < a.nz) {
    const int ci = cube_index(a, p);
    const int nt = a.count[ci];
    int64_t t = p ? tincl[p - 1] : 0;
    for (int q = 0; q < nt; ++q, ++t) {
      for (int r = 0; r < 3; ++r) {
        const int e = a.edges[ci * 3 * kMcMaxTris + 3 * q + r];
        const int ax = e >> 2, o0 = ax == 0 ? 1 : 0, o1 = ax == 2 ? 1 : 2;
        int off[3] = {0, 0, 0};
        off[o0] = e & 1;
        off[o1] = e >> 1 & 1;
        const int qi = i + off[0], qj = j + off[1], qk = k + off[2];
        const int64_t qp = (qi * sx + qj * sy) + qk;
        const int qb = owned_bits(a, qi, qj, qk, qp);
        faces[3 * t + r] = (qp ? vincl[qp - 1] : 0) + __popc(qb & ((1 << ax) - 1));
      }
    }
  }
}

}  // namespace vx

using namespace vx;

extern "C" {

int vx_mc_tables(int8_t* count, int8_t* edges) {
  VX_CHECK_ARG(count && edges, "mc_tables: missing output");
  const McTables& t = tables();
  memcpy(count, t.count, sizeof(t.count));
  memcpy(edges, t.edges, sizeof(t.edges));
  return VX_OK;
}

int vx_mc_max_tris(void) { return kMcMaxTris; }

int vx_extract_fields(const float* sdf, const int64_t* geom, const float* box, const float* bounds,
                      int n0, int n1, int n2, float scale, float* u, void* stream) {
  VX_CHECK_ARG(sdf && geom && box && bounds && u, "extract_fields: missing argument");
  VX_CHECK_ARG(geom[0] == 1, "extract_fields: the SDF grid must have one channel");
  VX_CHECK_ARG(n0 >= 1 && n1 >= 1 && n2 >= 1 && n0 <= 65535 && n1 <= 65535,
               "extract_fields: bad lattice size");
  FieldArgs a;
  a.sdf = sdf;
  a.g.C = 1;
  a.g.X = (int)geom[1];
  a.g.Y = (int)geom[2];
  a.g.Z = (int)geom[3];
  a.g.sc = geom[4];
  a.g.sx = geom[5];
  a.g.sy = geom[6];
  a.g.sz = geom[7];
  for (int i = 0; i < 3; ++i) {
    a.mn[i] = box[i];
    a.mx[i] = box[3 + i];
    a.bmin[i] = bounds[i];
    a.bmax[i] = bounds[3 + i];
  }
  a.n[0] = n0;
  a.n[1] = n1;
  a.n[2] = n2;
  a.scale = scale;
  a.u = u;
  const dim3 grid((unsigned)cdiv(n2, 256), (unsigned)n1, (unsigned)n0);
  k_extract_fields<<<grid, 256, 0, (hipStream_t)stream>>>(a);
  return launch_status("extract_fields");
}

// count: device copy of the [256] table counts, edges: [256 * 3 * vx_mc_max_tris()].
int vx_mc_count(const float* u, int nx, int ny, int nz, float iso, const int8_t* count,
                const int8_t* edges, int64_t* vcnt, int64_t* tcnt, void* stream) {
  VX_CHECK_ARG(u && count && edges && vcnt && tcnt, "mc_count: missing argument");
  VX_CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1 && nx <= 65535 && ny <= 65535, "mc_count: bad size");
  McArgs a{u, nx, ny, nz, iso, count, edges};
  const dim3 grid((unsigned)cdiv(nz, 256), (unsigned)ny, (unsigned)nx);
  k_mc_count<<<grid, 256, 0, (hipStream_t)stream>>>(a, vcnt, tcnt);
  return launch_status("mc_count");
}

int vx_mc_emit(const float* u, int nx, int ny, int nz, float iso, const int8_t* count,
               const int8_t* edges, const int64_t* vincl, const int64_t* tincl,
               const double* bounds, double* verts, int64_t* faces, void* stream) {
  VX_CHECK_ARG(u && count && edges && vincl && tincl, "mc_emit: missing argument");
  VX_CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1 && nx <= 65535 && ny <= 65535, "mc_emit: bad size");
  McArgs a{u, nx, ny, nz, iso, count, edges};
  double b[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  // float32 bounds, their difference taken in float32 (lib/dvgo_ori.py:699-702 scales
  // float64 vertices by float32 numpy arrays)
  if (bounds)
    for (int i = 0; i < 3; ++i) {
      b[i] = (double)(float)bounds[i];
      e[i] = (double)((float)bounds[3 + i] - (float)bounds[i]);
    }
  const dim3 grid((unsigned)cdiv(nz, 256), (unsigned)ny, (unsigned)nx);
  k_mc_emit<<<grid, 256, 0, (hipStream_t)stream>>>(a, vincl, tincl, bounds ? 1 : 0, b[0], b[1],
                                                   b[2], e[0], e[1], e[2], verts, faces);
  return launch_status("mc_emit");
}

}  // extern "C"

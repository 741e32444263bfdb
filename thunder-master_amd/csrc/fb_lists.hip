// Neighbour and common-neighbour (triplet) lists on the GPU, bit-identical to
// c.SYSTEM/M_neighbors.f90 (find_neighbors :356-519, find_common_neighbors
// :681-812).
//
// One workgroup per central atom walks the candidates in the reference's order
// (image mbeta outer, atom jatom inner; or ineigh outer, jneigh inner for the
// triplets) in chunks of the workgroup size; an ordered compaction (wave
// ballot + cross-wave prefix in LDS) appends the accepted candidates, so no
// sort is needed and the list order is the reference's.  Distances use the
// reference's expression with contraction off.  Images whose bounding box lies
// farther than the cutoff (plus a 1e-6 A margin) from the central atom are
// skipped: they cannot contain a neighbour.
#include "fb_device.h"

#define LWG 256
#include "fb_lists_dev.h"

// atoms binned by cell (counting sort of the neighbour search)
__global__ void k_cell_count(DView v, int *cnt, int *atom_cell) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= v.natoms) return;
  const DStruct S = v.st[v.atom_struct[ia]];
  const double r[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  const int c = S.cell_off + cell_of(S, r);
  atom_cell[ia] = c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_cell_fill(DView v, const int *atom_cell, int *cursor, int *cell_atoms) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= v.natoms) return;
  const int c = atom_cell[ia];
  const int k = atomicAdd(&cursor[c], 1);
  cell_atoms[v.cell_start[c] + k] = ia - v.st[v.atom_struct[ia]].atom_off;
}

// FILL = false: count neighbours (find_neigh_max); FILL = true: write them.
// With a cell list (v.cell_bits > 0) the candidates of each image are the
// atoms of the cells around r1 - xl(mbeta); the accepted ones are marked in a
// bitmap over jatom and written in jatom order, so the list order is the
// reference's either way.  Dynamic LDS: v.cell_bits words.
template <bool FILL>
__global__ __launch_bounds__(LWG) void k_neighbors(DView v, int *neighn_out, const int64_t *noff,
                                                   int *neigh_j, int *neigh_b, int *neigh_self,
                                                   int *pair_atom, int *status) {
  extern __shared__ unsigned s_bits[];
  __shared__ int64_t s_cs[27];
  __shared__ int s_pre[28];
  __shared__ int s_wave[LWG / 64];
  __shared__ int s_img[FB_NIMG];
  const int ia = blockIdx.x;  // global atom
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const int li = ia - S.atom_off;  // local 0-based
  const int in1 = v.imass[ia];
  double r1[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  const double rci = v.sp.rcmax[in1 - 1];
  const double reach = rci + S.rcmax_all + 1.0e-6;
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  int count = 0;
  int64_t base = FILL ? noff[ia] : 0;
  if (v.cell_bits > 0) {
    const int nw = (S.natoms + 31) >> 5;
    const int nimg = hit_images(S, xls, r1, reach, s_img, s_wave);
    for (int ki = 0; ki < nimg; ki++) {
      const int mb = s_img[ki];
      const double *xlb = xls + 3 * mb;
      const double q[3] = {r1[0] - xlb[0], r1[1] - xlb[1], r1[2] - xlb[2]};
      const int nc = query_cells(v, S, q, s_cs, s_pre);
      if (nc == 0) continue;
      for (int w = threadIdx.x; w < nw; w += LWG) s_bits[w] = 0u;
      __syncthreads();
      for (int t = threadIdx.x; t < s_pre[nc]; t += LWG) {
        const int lj = cell_candidate(v, s_cs, s_pre, nc, t);
        const int ja = S.atom_off + lj;
        double r2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) r2[k] = v.ratom[3 * ja + k] + xlb[k];
        const double z = d_distance(r1, r2);
        const double range_max = rci + v.sp.rcmax[v.imass[ja] - 1] - 0.01;
        if (lj != li && z < 0.5) atomicOr(status, ST_TOO_CLOSE);
        if (z < range_max) atomicOr(&s_bits[lj >> 5], 1u << (lj & 31));
      }
      __syncthreads();
      for (int w0 = 0; w0 < nw; w0 += LWG) {
        const int w = w0 + threadIdx.x;
        const unsigned bits = (w < nw) ? s_bits[w] : 0u;
        int tot;
        const int pos = block_scan_count(__popc(bits), s_wave, &tot);
        if (FILL) {
          unsigned b = bits;
          int r = 0;
          while (b) {
            const int lj = w * 32 + __ffs(b) - 1;
            b &= b - 1u;
            const int64_t p = base + count + pos + r;
            neigh_j[p] = lj + 1;
            neigh_b[p] = mb;
            pair_atom[p] = ia;
            if (lj == li && mb == 0) neigh_self[ia] = count + pos + r + 1;
            r++;
          }
        }
        count += tot;
      }
    }
    if (!FILL && threadIdx.x == 0) neighn_out[ia] = count;
    return;
  }
  const int nimg = hit_images(S, xls, r1, reach, s_img, s_wave);
  for (int ki = 0; ki < nimg; ki++) {
    const int mb = s_img[ki];
    const double *xlb = xls + 3 * mb;
    for (int j0 = 0; j0 < S.natoms; j0 += LWG) {
      int lj = j0 + threadIdx.x;
      bool pass = false;
      if (lj < S.natoms) {
        int ja = S.atom_off + lj;
        double r2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) r2[k] = v.ratom[3 * ja + k] + xlb[k];
        double z = d_distance(r1, r2);
        double range_max = rci + v.sp.rcmax[v.imass[ja] - 1] - 0.01;
        if (lj != li && z < 0.5) atomicOr(status, ST_TOO_CLOSE);
        pass = z < range_max;
      }
      int tot;
      int pos = block_rank(pass, s_wave, &tot);
      if (FILL && pass) {
        int64_t p = base + count + pos;
        neigh_j[p] = lj + 1;
        neigh_b[p] = mb;
        pair_atom[p] = ia;
        if (lj == li && mb == 0) neigh_self[ia] = count + pos + 1;
      }
      count += tot;
    }
  }
  if (!FILL && threadIdx.x == 0) neighn_out[ia] = count;
}

// displacement of every neighbour pair, xl(jbeta) + r_j - r_i with the
// reference's evaluation order (M_neighbors.f90:766-769 compares it with r2 - r1)
__global__ void k_pair_disp(DView v, double *disp) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.npairs) return;
  const int ia = v.pair_atom[p];
  const int s = v.atom_struct[ia];
  const int ja = v.st[s].atom_off + v.neigh_j[p] - 1;
  const double *xlb = v.xl + (int64_t)s * FB_NIMG * 3 + 3 * v.neigh_b[p];
#pragma unroll
  for (int k = 0; k < 3; k++) disp[3 * p + k] = xlb[k] + v.ratom[3 * ja + k] - v.ratom[3 * ia + k];
}

// neigh_back, c.SYSTEM/M_neighbors.f90:469-486: the jneigh of jatom that
// points back at (iatom, -xl(mbeta)) within 1e-3 A per component.
__global__ void k_neigh_back(DView v, int *neigh_back, int *status) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.npairs) return;
  int ia = v.pair_atom[p];
  int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  int li1 = ia - S.atom_off + 1;
  int ja = S.atom_off + v.neigh_j[p] - 1;
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const double *xm = xls + 3 * v.neigh_b[p];
  int found = -999;
  int64_t jo = v.noff[ja];
  int nj = v.neighn[ja];
  for (int jn = 0; jn < nj; jn++) {
    int ka = v.neigh_j[jo + jn];
    const double *xj = xls + 3 * v.neigh_b[jo + jn];
    if (li1 != ka || fabs(xm[0] + xj[0]) > 1.0e-3 || fabs(xm[1] + xj[1]) > 1.0e-3 ||
        fabs(xm[2] + xj[2]) > 1.0e-3) {
    } else {
      if (found != -999) atomicOr(status, ST_BACK);
      found = jn + 1;
    }
  }
  if (found == -999) atomicOr(status, ST_BACK);
  neigh_back[p] = found;
}

// common neighbours of ialpha.  FILL = false counts (find_common_max plus the
// drop rule of find_common_neighbors, so the counts are exact); FILL = true
// writes the triplet and its target pair.
template <bool FILL>
__global__ __launch_bounds__(LWG) void k_common(DView v, int *ncommon_out, const int64_t *toff,
                                                int4 *trip, int *trip_pair, int *status) {
  __shared__ int s_wave[LWG / 64];
  const int ial = blockIdx.x;
  const int s = v.atom_struct[ial];
  const DStruct S = v.st[s];
  const int lal = ial - S.atom_off + 1;  // 1-based local
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int64_t o = v.noff[ial];
  const int n = v.neighn[ial];
  const int64_t ncand = (int64_t)n * n;
  int count = 0;
  int64_t base = FILL ? toff[ial] : 0;
  for (int64_t c0 = 0; c0 < ncand; c0 += LWG) {
    int64_t c = c0 + threadIdx.x;
    bool pass = false;
    int ia = 0, ib = 0, ja = 0, jb = 0, m = -9999;
    if (c < ncand) {
      int in = (int)(c / n), jn = (int)(c % n);
      ia = v.neigh_j[o + in];
      ib = v.neigh_b[o + in];
      ja = v.neigh_j[o + jn];
      jb = v.neigh_b[o + jn];
      if (!(ia == lal && ib == 0) && !(ja == lal && jb == 0) && in != jn) {
        int gia = S.atom_off + ia - 1, gja = S.atom_off + ja - 1;
        double r1[3], r2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
          r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
        }
        double z = d_distance(r1, r2);
        double range_max = v.sp.rcmax[v.imass[gia] - 1] + v.sp.rcmax[v.imass[gja] - 1] - 0.01;
        if (z < range_max) {
          // mneigh: the first kneigh of iatom whose displacement r3 = xl(kbeta)
          // + r_k - r_i matches r2 - r1 within 1e-4 per component; eight
          // displacements loaded at a time, tested in order
          const double *dp = v.pair_disp + 3 * v.noff[gia];
          const int ni = v.neighn[gia];
          for (int k0 = 0; k0 < ni && m == -9999; k0 += 8) {
            double r3[8][3];
#pragma unroll
            for (int u = 0; u < 8; u++)
              if (k0 + u < ni)
#pragma unroll
                for (int k = 0; k < 3; k++) r3[u][k] = dp[3 * (k0 + u) + k];
#pragma unroll
            for (int u = 0; u < 8; u++)
              if (m == -9999 && k0 + u < ni && fabs(r3[u][0] - r2[0] + r1[0]) < 1.0e-4 &&
                  fabs(r3[u][1] - r2[1] + r1[1]) < 1.0e-4 && fabs(r3[u][2] - r2[2] + r1[2]) < 1.0e-4)
                m = k0 + u + 1;
          }
          pass = (m != -9999);
        }
      }
    }
    int tot;
    int pos = block_rank(pass, s_wave, &tot);
    if (FILL && pass) {
      int64_t t = base + count + pos;
      int gia = S.atom_off + ia - 1, gja = S.atom_off + ja - 1;
      trip[t] = make_int4(ial, gia, gja, (ib & 0xffff) | (jb << 16));
      trip_pair[t] = (int)(v.noff[gia] + m - 1);
      // mneigh is recovered from the target pair: m = trip_pair - noff[iatom] + 1
    }
    count += tot;
  }
  if (!FILL && threadIdx.x == 0) ncommon_out[ial] = count;
}

// segment bookkeeping after the stable sort of triplets by target pair
__global__ void k_seg_mark(int64_t ntrip, const int *sorted_pair, int *flag) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip) return;
  flag[t] = (t == 0 || sorted_pair[t] != sorted_pair[t - 1]) ? 1 : 0;
}

__global__ void k_seg_fill(int64_t ntrip, const int *sorted_pair, const int *segid_incl, int64_t *seg_off,
                           int *seg_pair, int *trip_seg, int *pair_tgt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip) return;
  int sid = segid_incl[t] - 1;
  trip_seg[t] = sid;
  if (t == 0 || sorted_pair[t] != sorted_pair[t - 1]) {
    seg_off[sid] = t;
    seg_pair[sid] = sorted_pair[t];
    pair_tgt[sorted_pair[t]] = 1;
  }
  if (t == ntrip - 1) seg_off[sid + 1] = ntrip;
}

__global__ void k_gather_trip(int64_t ntrip, const int *perm, const int4 *trip_in, int4 *trip_out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip) return;
  trip_out[t] = trip_in[perm[t]];
}

__global__ void k_iota(int64_t n, int *a) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = (int)t;
}

// block offsets: boff[p] = sum of norb_i * norb_j over earlier pairs; and
// the pair's record (the atoms, the image and the species in one int4)
__global__ void k_block_sizes(DView v, int64_t *sz, int4 *rec) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.npairs) return;
  int ia = v.pair_atom[p];
  int s = v.atom_struct[ia];
  int ja = v.st[s].atom_off + v.neigh_j[p] - 1;
  const int in1 = v.imass[ia], in2 = v.imass[ja];
  sz[p] = (int64_t)v.sp.norb[in1 - 1] * v.sp.norb[in2 - 1];
  rec[p] = make_int4(ia, ja, s * FB_NIMG + v.neigh_b[p], in1 | (in2 << 16));
}

// Tiles of whole target segments for the segment-owning kernels, packed per
// atom (a segment's target pair belongs to one atom, so a tile never needs to
// span atoms): greedily at most ch triplets per tile, a longer segment alone.
// One thread per atom over its segments, found by binary search of its pair
// range in the sorted segment pairs; FILL = false counts, FILL = true writes
// (first, last segment, first, end triplet).  Any partition into whole
// segments gives the kernels the same per-segment order, hence the same bits.
template <bool FILL>
__global__ void k_tiles_atom(int natoms, const int64_t *pair_off, const int64_t *seg_off, const int *seg_pair,
                             int64_t nseg, int ch, int *cnt, const int64_t *tile_off, int4 *tiles) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= natoms) return;
  auto lower = [&](int64_t key) {
    int64_t lo = 0, hi = nseg;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)seg_pair[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int64_t s0 = lower(pair_off[ia]), s1 = lower(pair_off[ia + 1]);
  int count = 0;
  for (int64_t s = s0; s < s1;) {
    int64_t e = s;
    while (e + 1 < s1 && seg_off[e + 2] - seg_off[s] <= ch) e++;
    if (FILL) tiles[tile_off[ia] + count] = make_int4((int)s, (int)e, (int)seg_off[s], (int)seg_off[e + 1]);
    count++;
    s = e + 1;
  }
  if (!FILL) cnt[ia] = count;
}

__global__ void k_gather_i64(int n, const int64_t *src, const int64_t *idx, int64_t *dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[idx[k]];
}

// ----------------------------------------------------------------- launchers
#include <hipcub/hipcub.hpp>

extern "C" {

int fbk_tiles_atom(int natoms, const int64_t *pair_off, const int64_t *seg_off, const int *seg_pair, int64_t nseg,
                   int ch, int fill, int *cnt, const int64_t *tile_off, int4 *tiles, hipStream_t st) {
  if (natoms == 0) return 0;
  const dim3 g((unsigned)((natoms + 255) / 256));
  if (fill)
    hipLaunchKernelGGL(k_tiles_atom<true>, g, dim3(256), 0, st, natoms, pair_off, seg_off, seg_pair, nseg, ch, cnt,
                       tile_off, tiles);
  else
    hipLaunchKernelGGL(k_tiles_atom<false>, g, dim3(256), 0, st, natoms, pair_off, seg_off, seg_pair, nseg, ch, cnt,
                       tile_off, tiles);
  return (int)hipGetLastError();
}

int fbk_gather_i64(int n, const int64_t *src, const int64_t *idx, int64_t *dst, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_gather_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src, idx, dst);
  return (int)hipGetLastError();
}

int fbk_cell_count(const DView &v, int *cnt, int *atom_cell, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_cell_count, dim3((unsigned)((v.natoms + 255) / 256)), dim3(256), 0, st, v, cnt, atom_cell);
  return (int)hipGetLastError();
}

int fbk_cell_fill(const DView &v, const int *atom_cell, int *cursor, int *cell_atoms, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_cell_fill, dim3((unsigned)((v.natoms + 255) / 256)), dim3(256), 0, st, v, atom_cell, cursor,
                     cell_atoms);
  return (int)hipGetLastError();
}

int fbk_neighbors_count(const DView &v, int *neighn, int *status, hipStream_t st) {
  const size_t lds = sizeof(unsigned) * (size_t)v.cell_bits;
  hipLaunchKernelGGL(k_neighbors<false>, dim3(v.natoms), dim3(LWG), lds, st, v, neighn, (const int64_t *)nullptr,
                     (int *)nullptr, (int *)nullptr, (int *)nullptr, (int *)nullptr, status);
  return (int)hipGetLastError();
}

int fbk_neighbors_fill(const DView &v, const int64_t *noff, int *neigh_j, int *neigh_b, int *neigh_self,
                       int *pair_atom, int *status, hipStream_t st) {
  const size_t lds = sizeof(unsigned) * (size_t)v.cell_bits;
  hipLaunchKernelGGL(k_neighbors<true>, dim3(v.natoms), dim3(LWG), lds, st, v, (int *)nullptr, noff, neigh_j,
                     neigh_b, neigh_self, pair_atom, status);
  return (int)hipGetLastError();
}

int fbk_pair_disp(const DView &v, double *disp, hipStream_t st) {
  if (v.npairs == 0) return 0;
  hipLaunchKernelGGL(k_pair_disp, dim3((unsigned)((v.npairs + 255) / 256)), dim3(256), 0, st, v, disp);
  return (int)hipGetLastError();
}

int fbk_neigh_back(const DView &v, int *neigh_back, int *status, hipStream_t st) {
  if (v.npairs == 0) return 0;
  hipLaunchKernelGGL(k_neigh_back, dim3((unsigned)((v.npairs + 255) / 256)), dim3(256), 0, st, v, neigh_back,
                     status);
  return (int)hipGetLastError();
}

int fbk_common_count(const DView &v, int *ncommon, int *status, hipStream_t st) {
  hipLaunchKernelGGL(k_common<false>, dim3(v.natoms), dim3(LWG), 0, st, v, ncommon, (const int64_t *)nullptr,
                     (int4 *)nullptr, (int *)nullptr, status);
  return (int)hipGetLastError();
}

int fbk_common_fill(const DView &v, const int64_t *toff, int4 *trip, int *trip_pair, int *status,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_common<true>, dim3(v.natoms), dim3(LWG), 0, st, v, (int *)nullptr, toff, trip, trip_pair,
                     status);
  return (int)hipGetLastError();
}

// Lane order of the 2c workgroups (fb_2c_sp.h): within each 64-pair
// wavefront group of a workgroup item (the groups that pass c0 of item b
// forms: c0 + 64 w), the pairs by species pair, then bond length, then index,
// so that neighbouring lanes interpolate nearby rows of the same table.  perm[slot] is the
// group-relative index of the pair that the lane at slot takes.  Blocks and
// sums are placed by pair index, so any permutation gives the same bits.
__global__ __launch_bounds__(FB_2C_PCAP) void k_pair_perm(DView v, unsigned char *perm) {
  __shared__ double s_x[FB_2C_PCAP];
  __shared__ int s_k[FB_2C_PCAP];
  const int tid = threadIdx.x;
  const int64_t pb = v.noff[v.wg2c[blockIdx.x]], pe = v.noff[v.wg2c[blockIdx.x + 1]];
  for (int64_t c0 = pb; c0 < pe; c0 += FB_2C_PCAP) {
    const int64_t slot = c0 + tid;
    double x = 0.0;
    int key = 0;
    if (slot < pe) {
      const int4 pr = v.pair_rec[slot];
      key = pr.w;
      const double *xlp = v.xl + (int64_t)pr.z * 3;
      for (int q = 0; q < 3; q++) {
        const double d = v.ratom[3 * pr.y + q] + xlp[q] - v.ratom[3 * pr.x + q];
        x = x + d * d;
      }
    }
    __syncthreads();   // the previous pass has read s_x
    s_x[tid] = x;
    s_k[tid] = key;
    __syncthreads();
    if (slot < pe) {
      const int g0 = tid & ~63;
      const int n = (int)min((int64_t)64, pe - (c0 + g0));
      int rank = 0;
      for (int j = 0; j < n; j++) {
        const double xj = s_x[g0 + j];
        const int kj = s_k[g0 + j];
        rank += (kj < key || (kj == key && (xj < x || (xj == x && g0 + j < tid)))) ? 1 : 0;
      }
      perm[c0 + g0 + rank] = (unsigned char)(tid & 63);
    }
  }
}

int fbk_pair_perm(const DView &v, unsigned char *perm, hipStream_t st) {
  if (v.nwg2c == 0) return 0;
  hipLaunchKernelGGL(k_pair_perm, dim3((unsigned)v.nwg2c), dim3(FB_2C_PCAP), 0, st, v, perm);
  return (int)hipGetLastError();
}

int fbk_block_sizes(const DView &v, int64_t *sz, int4 *rec, hipStream_t st) {
  if (v.npairs == 0) return 0;
  hipLaunchKernelGGL(k_block_sizes, dim3((unsigned)((v.npairs + 255) / 256)), dim3(256), 0, st, v, sz, rec);
  return (int)hipGetLastError();
}

// exclusive prefix sums (int -> int64 and int64 -> int64); out has n+1 entries
int fbk_scan_i32(const int *in, int64_t *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st) {
  // out[0] = 0, out[i+1] = sum(in[0..i])
  if (tmp == nullptr) {
    return (int)hipcub::DeviceScan::InclusiveSum(nullptr, *tmp_bytes, in, out + 1, (int)n, st);
  }
  hipError_t e = hipMemsetAsync(out, 0, sizeof(int64_t), st);
  if (e != hipSuccess) return (int)e;
  if (n == 0) return 0;
  return (int)hipcub::DeviceScan::InclusiveSum(tmp, *tmp_bytes, in, out + 1, (int)n, st);
}

int fbk_scan_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st) {
  if (tmp == nullptr) {
    return (int)hipcub::DeviceScan::InclusiveSum(nullptr, *tmp_bytes, in, out + 1, (int)n, st);
  }
  hipError_t e = hipMemsetAsync(out, 0, sizeof(int64_t), st);
  if (e != hipSuccess) return (int)e;
  if (n == 0) return 0;
  return (int)hipcub::DeviceScan::InclusiveSum(tmp, *tmp_bytes, in, out + 1, (int)n, st);
}

// stable sort of triplet ids by target pair; then segments
int fbk_sort_pairs(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int64_t n, int end_bit,
                   void *tmp, size_t *tmp_bytes, hipStream_t st) {
  return (int)hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, keys_in, keys_out, vals_in, vals_out, (int)n,
                                                 0, end_bit, st);
}

int fbk_scan_incl_i32(const int *in, int *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st) {
  return (int)hipcub::DeviceScan::InclusiveSum(tmp, *tmp_bytes, in, out, (int)n, st);
}

int fbk_iota(int *a, int64_t n, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a);
  return (int)hipGetLastError();
}

int fbk_seg_mark(int64_t ntrip, const int *sorted_pair, int *flag, hipStream_t st) {
  if (ntrip == 0) return 0;
  hipLaunchKernelGGL(k_seg_mark, dim3((unsigned)((ntrip + 255) / 256)), dim3(256), 0, st, ntrip, sorted_pair,
                     flag);
  return (int)hipGetLastError();
}

int fbk_seg_fill(int64_t ntrip, const int *sorted_pair, const int *segid_incl, int64_t *seg_off, int *seg_pair,
                 int *trip_seg, int *pair_tgt, hipStream_t st) {
  if (ntrip == 0) return 0;
  hipLaunchKernelGGL(k_seg_fill, dim3((unsigned)((ntrip + 255) / 256)), dim3(256), 0, st, ntrip, sorted_pair,
                     segid_incl, seg_off, seg_pair, trip_seg, pair_tgt);
  return (int)hipGetLastError();
}

int fbk_gather_trip(int64_t ntrip, const int *perm, const int4 *in, int4 *out, hipStream_t st) {
  if (ntrip == 0) return 0;
  hipLaunchKernelGGL(k_gather_trip, dim3((unsigned)((ntrip + 255) / 256)), dim3(256), 0, st, ntrip, perm, in,
                     out);
  return (int)hipGetLastError();
}

}  // extern "C"

__global__ void k_fill_i32(int64_t n, int *a, int val) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = val;
}

extern "C" int fbk_fill_i32(int *a, int64_t n, int val, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, val);
  return (int)hipGetLastError();
}

// species group (neutral bcna) and structure of every sorted triplet
__global__ void k_trip_aux(DView v, int *aux) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= v.ntrip) return;
  const int4 tr = v.trip[t];
  const int in1 = v.imass[tr.y], in2 = v.imass[tr.z], in3 = v.imass[tr.x];
  int gid = v.idx3c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * FB_MAXSUB3];
  aux[t] = (gid < 0 ? 0xffff : gid) | (v.atom_struct[tr.x] << 16);
}

// segment ownership of every tile of ch consecutive sorted triplets: the
// segments that start inside the tile, and the triplets they span
extern "C" int fbk_trip_aux(const DView &v, int *aux, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  hipLaunchKernelGGL(k_trip_aux, dim3((unsigned)((v.ntrip + 255) / 256)), dim3(256), 0, st, v, aux);
  return (int)hipGetLastError();
}


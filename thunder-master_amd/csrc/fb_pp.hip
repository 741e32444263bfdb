// Non-local pseudopotential path on the GPU:
//  * driver_neighbors_PP (c.SYSTEM/M_neighbors_PP.f90): the PP, PPx and PPp
//    neighbour lists with their maps and points, and the PP common-neighbour
//    triplets, bit-identical to the reference's order;
//  * assemble_svnl (j.ASSEMBLERS/M_assemble_PP_2c.f90:95-207): the rotated
//    <phi_mu(iatom) | Psi_ncc(jatom)> blocks of every PP pair;
//  * assemble_vnl_2c (:230-496) and assemble_vnl_3c (M_assemble_PP_3c.f90:94-237):
//    vnl(iatom, k) = sum over the PP atoms alpha near both of
//    sum_ncc cl(ncc) sVNL(i, alpha)(mu, ncc) sVNL(j, alpha)(nu, ncc), in the
//    reference's order (atom case, ontop left, ontop right, then the triplets in
//    list order), one thread per target block or per target segment.
// Binary64 throughout; contraction off; deterministic.
#include "fb_device.h"

#define LWG 256
#include "fb_lists_dev.h"
#include "fb_pp.h"

#define NBP 16    // elements of one block (at most 4 x 4)
#define T3P 64    // triplets per vnl_3c tile

// ------------------------------------------------------------------ lists
// PP (rc_i + rcutoff_PP_j) or PPx (rcutoff_PP_i + rc_j) neighbours, both with
// the reference's strict test z < range - 0.01 (M_neighbors_PP.f90:466-474,
// 730-738); FILL = false counts, FILL = true writes.
// With the cell list of the H+S search (when its cells are at least as wide as
// this list's reach) the candidates of each image come from the cells around
// r1 - xl(mbeta) and the accepted ones are written in jatom order through a
// bitmap, as in k_neighbors.  Dynamic LDS: v.cell_bits words.
template <bool FILL>
__global__ __launch_bounds__(LWG) void k_pp_neighbors(DView v, DPP q, int list) {
  extern __shared__ unsigned s_bits[];
  __shared__ int64_t s_cs[27];
  __shared__ int s_pre[28];
  __shared__ int s_wave[LWG / 64];
  __shared__ int s_img[FB_NIMG];
  const int ia = blockIdx.x;
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const int li = ia - S.atom_off;
  const int in1 = v.imass[ia];
  const double r1[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  const double rci = (list == PPL_PP) ? v.sp.rcmax[in1 - 1] : q.rcpp[in1 - 1];
  double maxb = -99.0;
  for (int k = 0; k < v.sp.ns; k++) maxb = fmax(maxb, (list == PPL_PP) ? q.rcpp[k] : v.sp.rcmax[k]);
  const double reach = rci + maxb + 1.0e-6;
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  int count = 0;
  const int64_t base = FILL ? q.off[list][ia] : 0;
  if (v.cell_bits > 0 && reach <= S.crange) {
    const int nw = (S.natoms + 31) >> 5;
    const int nimg = hit_images(S, xls, r1, reach, s_img, s_wave);
    for (int ki = 0; ki < nimg; ki++) {
      const int mb = s_img[ki];
      const double *xlb = xls + 3 * mb;
      const double qp[3] = {r1[0] - xlb[0], r1[1] - xlb[1], r1[2] - xlb[2]};
      const int nc = query_cells(v, S, qp, s_cs, s_pre);
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
        const double rcj = (list == PPL_PP) ? q.rcpp[v.imass[ja] - 1] : v.sp.rcmax[v.imass[ja] - 1];
        if (z < rci + rcj - 0.01) atomicOr(&s_bits[lj >> 5], 1u << (lj & 31));
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
            q.j[list][p] = lj + 1;
            q.b[list][p] = mb;
            q.atom[list][p] = ia;
            if (lj == li && mb == 0) q.self[list][ia] = count + pos + r + 1;
            r++;
          }
        }
        count += tot;
      }
    }
    if (!FILL && threadIdx.x == 0) q.n[list][ia] = count;
    return;
  }
  const int nimg = hit_images(S, xls, r1, reach, s_img, s_wave);
  for (int ki = 0; ki < nimg; ki++) {
    const int mb = s_img[ki];
    const double *xlb = xls + 3 * mb;
    for (int j0 = 0; j0 < S.natoms; j0 += LWG) {
      const int lj = j0 + threadIdx.x;
      bool pass = false;
      if (lj < S.natoms) {
        const int ja = S.atom_off + lj;
        double r2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) r2[k] = v.ratom[3 * ja + k] + xlb[k];
        const double z = d_distance(r1, r2);
        const double rcj = (list == PPL_PP) ? q.rcpp[v.imass[ja] - 1] : v.sp.rcmax[v.imass[ja] - 1];
        pass = z < rci + rcj - 0.01;
      }
      int tot;
      const int pos = block_rank(pass, s_wave, &tot);
      if (FILL && pass) {
        const int64_t p = base + count + pos;
        q.j[list][p] = lj + 1;
        q.b[list][p] = mb;
        q.atom[list][p] = ia;
        if (lj == li && mb == 0) q.self[list][ia] = count + pos + 1;
      }
      count += tot;
    }
  }
  if (!FILL && threadIdx.x == 0) q.n[list][ia] = count;
}

// make_cells order of the images (c.SYSTEM/make_cells.f90:130-144): iz outer,
// iy, ix inner over -4..4 without the origin, which is image 0
__device__ __forceinline__ void image_coords(int mb, int *c) {
  if (mb == 0) { c[0] = c[1] = c[2] = 0; return; }
  const int L = (mb <= 364) ? mb - 1 : mb;
  c[0] = L % 9 - 4;
  c[1] = (L / 9) % 9 - 4;
  c[2] = L / 81 - 4;
}

// the first image mbeta with |xl(mbeta) - (xl(jbeta) + xl(ibeta))| < 1e-4 in
// make_cells order, mbeta_max + 1 when none (find_neighbors_PPp,
// M_neighbors_PP.f90:983-990).  The integer sum of the two image coordinates
// names the only image that can match; it is checked with the reference's
// distance, and the full scan covers everything else.
__device__ inline int image_of_sum(const DStruct &S, const double *xls, int ib, int jb) {
  double r3[3];
#pragma unroll
  for (int k = 0; k < 3; k++) r3[k] = xls[3 * jb + k] + xls[3 * ib + k];
  if (S.mbeta_max == FB_NIMG - 1) {
    int a[3], c[3];
    image_coords(ib, a);
    image_coords(jb, c);
    const int x = a[0] + c[0], y = a[1] + c[1], z = a[2] + c[2];
    if (abs(x) <= 4 && abs(y) <= 4 && abs(z) <= 4) {
      int m = 0;
      if (x || y || z) {
        const int L = (z + 4) * 81 + (y + 4) * 9 + (x + 4);
        m = (L < 364) ? L + 1 : L;
      }
      // two images within 1e-4 of one point would need a lattice vector below 2e-4 A
      if (d_distance(xls + 3 * m, r3) < 1.0e-4) return m;
    }
  }
  for (int mb = 0; mb <= S.mbeta_max; mb++)
    if (d_distance(xls + 3 * mb, r3) < 1.0e-4) return mb;
  return S.mbeta_max + 1;
}

// candidates of the PPp list of each atom: sum over its PP neighbours of their PPx counts
__global__ void k_ppp_candidates(DView v, DPP q, int64_t *cnt) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= v.natoms) return;
  const int off = v.st[v.atom_struct[ia]].atom_off;
  int64_t c = 0;
  for (int64_t p = q.off[PPL_PP][ia]; p < q.off[PPL_PP][ia + 1]; p++) c += q.n[PPL_PPX][off + q.j[PPL_PP][p] - 1];
  cnt[ia] = c;
}

// neighbors_PPp (M_neighbors_PP.f90:922-1012), one workgroup per atom.  The
// candidates (ineigh over PP, jneigh over the PPx list of that neighbour, in
// this order) get their key (jatom, image of ibeta + jbeta) in the atom's
// scratch range (k_ppp_keys); a candidate is kept when no earlier candidate has
// the same key, which is the reference's first-seen rule, and the kept ones
// are compacted in candidate order (k_ppp_build: FILL = false counts, FILL =
// true writes the list and the self entry).
__global__ __launch_bounds__(LWG) void k_ppp_keys(DView v, DPP q) {
  const int ia = blockIdx.x;
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const int64_t c0 = q.cand_off[ia];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  int64_t base = 0;
  for (int64_t p = q.off[PPL_PP][ia]; p < q.off[PPL_PP][ia + 1]; p++) {
    const int ial = S.atom_off + q.j[PPL_PP][p] - 1;
    const int ib = q.b[PPL_PP][p];
    const int64_t r0 = q.off[PPL_PPX][ial], nr = q.off[PPL_PPX][ial + 1] - r0;
    for (int64_t r = threadIdx.x; r < nr; r += LWG) {
      q.cand_j[c0 + base + r] = q.j[PPL_PPX][r0 + r];
      q.cand_b[c0 + base + r] = image_of_sum(S, xls, ib, q.b[PPL_PPX][r0 + r]);
    }
    base += nr;
  }
}

// First occurrences by key through an LDS hash table: every candidate inserts
// its key (linear probing) and lowers the key's smallest candidate index; a
// candidate is kept when it holds that minimum.  A workgroup whose keys
// overflow the table, or whose candidates or atoms exceed its encoding, falls
// back to comparing each candidate with all earlier ones.
#define PPH 2048   // hash slots per workgroup

__device__ __forceinline__ unsigned ppp_hash(unsigned key) { return (key * 2654435761u) >> (32 - 11); }

template <bool FILL>
__global__ __launch_bounds__(LWG) void k_ppp_build(DView v, DPP q, int *status) {
  __shared__ int s_wave[LWG / 64];
  __shared__ int s_self;
  __shared__ unsigned s_key[PPH];
  __shared__ int s_min[PPH];
  __shared__ int s_over;
  const int ia = blockIdx.x;
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const int64_t c0 = q.cand_off[ia];
  const int64_t nc = q.cand_off[ia + 1] - c0;
  const int li = ia - S.atom_off + 1;
  const int64_t o = FILL ? q.off[PPL_PPP][ia] : 0;
  if (FILL && threadIdx.x == 0) s_self = -999;
  // key = jatom * 1024 + image (images 0..729): unique below 2^22 atoms
  const bool hashed = S.natoms < (1 << 22) && nc < (1ll << 31);
  if (hashed) {
    for (int k = threadIdx.x; k < PPH; k += LWG) {
      s_key[k] = 0xffffffffu;
      s_min[k] = 0x7fffffff;
    }
    if (threadIdx.x == 0) s_over = 0;
    __syncthreads();
    for (int64_t c = threadIdx.x; c < nc; c += LWG) {
      const unsigned key = (unsigned)q.cand_j[c0 + c] * 1024u + (unsigned)q.cand_b[c0 + c];
      unsigned h = ppp_hash(key);
      int probe = 0;
      while (probe < PPH) {
        const unsigned old = atomicCAS(&s_key[h], 0xffffffffu, key);
        if (old == 0xffffffffu || old == key) break;
        h = (h + 1) & (PPH - 1);
        probe++;
      }
      if (probe == PPH) s_over = 1;
      else atomicMin(&s_min[h], (int)c);
    }
    __syncthreads();
  }
  const bool use_hash = hashed && s_over == 0;
  int count = 0;
  for (int64_t k0 = 0; k0 < nc; k0 += LWG) {
    const int64_t c = k0 + threadIdx.x;
    bool keep = false;
    int j = 0, b = 0;
    if (c < nc) {
      j = q.cand_j[c0 + c];
      b = q.cand_b[c0 + c];
      bool found = false;
      if (use_hash) {
        // every key was inserted above; the probe is bounded all the same
        const unsigned key = (unsigned)j * 1024u + (unsigned)b;
        unsigned h = ppp_hash(key);
        for (int probe = 0; probe < PPH; probe++) {
          if (s_key[h] == key) {
            keep = s_min[h] == (int)c;
            found = true;
            break;
          }
          h = (h + 1) & (PPH - 1);
        }
      }
      if (!found) {
        keep = true;
        for (int64_t e = 0; e < c && keep; e++) keep = !(q.cand_j[c0 + e] == j && q.cand_b[c0 + e] == b);
      }
    }
    int tot;
    const int pos = block_rank(keep, s_wave, &tot);
    if (FILL && keep) {
      q.j[PPL_PPP][o + count + pos] = j;
      q.b[PPL_PPP][o + count + pos] = b;
      q.atom[PPL_PPP][o + count + pos] = ia;
      if (j == li && b == 0) s_self = count + pos + 1;   // a key occurs once in the kept list
    }
    count += tot;
  }
  if (!FILL) {
    if (threadIdx.x == 0) q.n[PPL_PPP][ia] = count;
    return;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    q.self[PPL_PPP][ia] = s_self;
    if (s_self == -999) atomicOr(status, ST_SELF);
  }
}

// maps of the PP and PPx pairs into PPp (first equal entry, M_neighbors_PP.f90:1026-1068)
// and the PPx points (index of iatom among jatom's PP neighbours, :1071-1088)
__global__ void k_pp_maps(DView v, DPP q, int *status) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t npp = q.npairs[PPL_PP];
  if (t >= npp + q.npairs[PPL_PPX]) return;
  const int list = (t < npp) ? PPL_PP : PPL_PPX;
  const int64_t p = (t < npp) ? t : t - npp;
  const int ia = q.atom[list][p];
  const int j = q.j[list][p], b = q.b[list][p];
  const int64_t o2 = q.off[PPL_PPP][ia];
  const int n2 = q.n[PPL_PPP][ia];
  int m = 0;
  for (int k = 0; k < n2; k++)
    if (q.j[PPL_PPP][o2 + k] == j && q.b[PPL_PPP][o2 + k] == b) { m = k + 1; break; }
  if (m == 0) atomicOr(status, ST_BACK);
  if (list == PPL_PP) {
    q.pp_map[p] = m;
    if (m) q.inv_pp[o2 + m - 1] = (int)p;
    return;
  }
  q.px_map[p] = m;
  if (m) q.inv_px[o2 + m - 1] = (int)p;
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int ja = S.atom_off + j - 1;
  const double r1[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  int point = 0;
  for (int64_t k = q.off[PPL_PP][ja]; k < q.off[PPL_PP][ja + 1]; k++) {
    const int ka = S.atom_off + q.j[PPL_PP][k] - 1, kb = q.b[PPL_PP][k];
    double r2[3];
#pragma unroll
    for (int c = 0; c < 3; c++) r2[c] = v.ratom[3 * ka + c] + xls[3 * kb + c] + xls[3 * b + c];
    if (d_distance(r1, r2) < 1.0e-4) { point = (int)(k - q.off[PPL_PP][ja]) + 1; break; }
  }
  if (point == 0) atomicOr(status, ST_BACK);
  q.px_point[p] = point;
}

// displacement of every PPp pair, xl(kbeta) + r_k - r_i in the reference's
// order (M_neighbors_PP.f90 common-neighbour match), NaN beyond the images
__global__ void k_ppp_disp(DView v, DPP q) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.npairs[PPL_PPP]) return;
  const int ia = q.atom[PPL_PPP][p];
  const int s = v.atom_struct[ia];
  const DStruct S = v.st[s];
  const int kb = q.b[PPL_PPP][p];
  double *d = q.ppp_disp + 3 * p;
  if (kb > S.mbeta_max) {
    d[0] = d[1] = d[2] = __builtin_nan("");
    return;
  }
  const int ka = S.atom_off + q.j[PPL_PPP][p] - 1;
  const double *xlb = v.xl + (int64_t)s * FB_NIMG * 3 + 3 * kb;
#pragma unroll
  for (int k = 0; k < 3; k++) d[k] = xlb[k] + v.ratom[3 * ka + k] - v.ratom[3 * ia + k];
}

__device__ __forceinline__ unsigned ppk_mix(unsigned k) {
  k ^= k >> 16;
  k *= 0x45d9f3bu;
  k ^= k >> 16;
  return k;
}

// slots of each atom's PPp key table: a power of two, at least twice the list
__global__ void k_ppk_size(DView v, DPP q, int *sz) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= v.natoms) return;
  const int n = q.n[PPL_PPP][ia];
  int t = 2;
  while (t < 2 * n) t <<= 1;
  sz[ia] = t;
}

// each PPp entry inserts its key into its atom's table (the keys of one list
// are distinct: the list keeps first occurrences)
__global__ void k_ppk_insert(DView v, DPP q) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.npairs[PPL_PPP]) return;
  const int ia = q.atom[PPL_PPP][p];
  const unsigned key = (unsigned)q.j[PPL_PPP][p] * 1024u + (unsigned)q.b[PPL_PPP][p];
  const int64_t o = q.pk_off[ia];
  const unsigned mask = (unsigned)(q.pk_off[ia + 1] - o - 1);
  unsigned h = ppk_mix(key) & mask;
  for (unsigned probe = 0; probe <= mask; probe++) {
    const unsigned old = atomicCAS(&q.pk_key[o + h], 0xffffffffu, key);
    if (old == 0xffffffffu) {
      q.pk_idx[o + h] = (int)(p - q.off[PPL_PPP][ia]);
      return;
    }
    if (old == key) return;
    h = (h + 1) & mask;
  }
}

// find_common_neighbors_PP (M_neighbors_PP.f90:1301-1420): ordered pairs of
// PPx neighbours of ialpha, neither ialpha itself, kept when the bond vector
// names an entry of iatom's PPp list (first within 1e-4).  FILL = false counts,
// FILL = true writes the triplet and its target PPp pair.
template <bool FILL>
__global__ __launch_bounds__(LWG) void k_pp_common(DView v, DPP q) {
  __shared__ int s_wave[LWG / 64];
  const int ial = blockIdx.x;
  const int s = v.atom_struct[ial];
  const DStruct S = v.st[s];
  const int lal = ial - S.atom_off + 1;
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int64_t o = q.off[PPL_PPX][ial];
  const int n = q.n[PPL_PPX][ial];
  const int64_t ncand = (int64_t)n * n;
  int count = 0;
  const int64_t base = FILL ? q.pc_off[ial] : 0;
  for (int64_t c0 = 0; c0 < ncand; c0 += LWG) {
    const int64_t c = c0 + threadIdx.x;
    bool pass = false;
    int ia = 0, ib = 0, ja = 0, jb = 0, m = -9999, gia = 0;
    if (c < ncand) {
      const int in = (int)(c / n), jn = (int)(c % n);
      ia = q.j[PPL_PPX][o + in];
      ib = q.b[PPL_PPX][o + in];
      ja = q.j[PPL_PPX][o + jn];
      jb = q.b[PPL_PPX][o + jn];
      if (!(ia == lal && ib == 0) && !(ja == lal && jb == 0) && in != jn) {
        gia = S.atom_off + ia - 1;
        const int gja = S.atom_off + ja - 1;
        double r1[3], r2[3], r[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
          r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
          r[k] = r2[k] - r1[k];
        }
        // first kneigh of iatom's PPp list whose displacement lies within
        // 1e-4 of r2 - r1 (entries with an image outside xl are NaN: never
        // a match, as the reference's search skips them).  The entry that
        // can match is (jatom, image with the coordinates of jbeta minus
        // ibeta): looked up by key and accepted with the reference's test.
        // No other entry can pass that test (two entries within 2e-4 A of
        // one point would be coincident atoms), so it is the first match.
        // Otherwise the list is scanned in order, eight at a time.
        const double *dp = q.ppp_disp + 3 * q.off[PPL_PPP][gia];
        const int ni = q.n[PPL_PPP][gia];
        {
          int a[3], c[3];
          image_coords(ib, a);
          image_coords(jb, c);
          const int x = c[0] - a[0], y = c[1] - a[1], z = c[2] - a[2];
          if (abs(x) <= 4 && abs(y) <= 4 && abs(z) <= 4 && (S.mbeta_max == FB_NIMG - 1 || (x | y | z) == 0)) {
            int mb = 0;
            if (x || y || z) {
              const int L = (z + 4) * 81 + (y + 4) * 9 + (x + 4);
              mb = (L < 364) ? L + 1 : L;
            }
            const unsigned key = (unsigned)ja * 1024u + (unsigned)mb;
            const int64_t to = q.pk_off[gia];
            const unsigned mask = (unsigned)(q.pk_off[gia + 1] - to - 1);
            unsigned h = ppk_mix(key) & mask;
            for (unsigned probe = 0; probe <= mask; probe++) {
              const unsigned kk = q.pk_key[to + h];
              if (kk == 0xffffffffu) break;
              if (kk == key) {
                const int k = q.pk_idx[to + h];
                if (d_distance(r, dp + 3 * k) < 1.0e-4) m = k + 1;
                break;
              }
              h = (h + 1) & mask;
            }
          }
        }
        for (int k0 = 0; k0 < ni && m == -9999; k0 += 8) {
          double r3[8][3];
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (k0 + u < ni)
#pragma unroll
              for (int k = 0; k < 3; k++) r3[u][k] = dp[3 * (k0 + u) + k];
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (m == -9999 && k0 + u < ni && d_distance(r, r3[u]) < 1.0e-4) m = k0 + u + 1;
        }
        pass = (m != -9999);
      }
    }
    int tot;
    const int pos = block_rank(pass, s_wave, &tot);
    if (FILL && pass) {
      const int64_t t = base + count + pos;
      const int gja = S.atom_off + ja - 1;
      q.trip[t] = make_int4(ial, gia, gja, (ib & 0xffff) | (jb << 16));
      q.tpair[t] = (int)(q.off[PPL_PPP][gia] + m - 1);
    }
    count += tot;
  }
  if (!FILL && threadIdx.x == 0) q.pc_n[ial] = count;
}

// block sizes: which 0 sVNL per PP pair (norb_mu x norb_PP_nu), 1 vnl per PPp pair
__global__ void k_pp_block_sizes(DView v, DPP q, int which, int64_t *sz) {
  const int list = which == 0 ? PPL_PP : PPL_PPP;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.npairs[list]) return;
  const int ia = q.atom[list][p];
  const int ja = v.st[v.atom_struct[ia]].atom_off + q.j[list][p] - 1;
  const int in2 = v.imass[ja];
  sz[p] = (int64_t)v.sp.norb[v.imass[ia] - 1] * (which == 0 ? q.norb_pp[in2 - 1] : v.sp.norb[in2 - 1]);
}

// ------------------------------------------------------------------ blocks
// rotate_PP (c.SYSTEM/rotate_PP.f90): orbital shells of in1 on the left, PP
// shells of in2 (s and p) on the right, the loop order and association of
// rotate; dm: the d rotation, read only for a d orbital shell
__device__ inline void d_rotate_pp(const DSpecies &sp, const DPP &q, int in1, int in2, const double *pm,
                                   const double *dm, const double *M, int nmu, double *X) {
  const int nnu = q.norb_pp[in2 - 1];
  for (int i = 0; i < nmu * nnu; i++) X[i] = 0.0;
  int n1 = 0;
  for (int issh = 0; issh < sp.nssh[in1 - 1]; issh++) {
    const int l1 = sp.lssh[in1 - 1][issh];
    int n2 = 0;
    for (int jssh = 0; jssh < q.nssh_pp[in2 - 1]; jssh++) {
      const int l2 = q.lssh_pp[in2 - 1][jssh];
      for (int m2 = 0; m2 < 2 * l2 + 1; m2++)
        for (int m1 = 0; m1 < 2 * l1 + 1; m1++)
          for (int k2 = 0; k2 < 2 * l2 + 1; k2++)
            for (int k1 = 0; k1 < 2 * l1 + 1; k1++) {
              const double left = (l1 == 0) ? 1.0 : (l1 == 1 ? pm[3 * m1 + k1] : dm[5 * m1 + k1]);
              const double right = (l2 == 0) ? 1.0 : pm[3 * m2 + k2];
              double *x = &X[(n2 + k2) * nmu + (n1 + k1)];
              *x = *x + left * M[(n2 + m2) * nmu + (n1 + m1)] * right;
            }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

// assemble_svnl: one lane per PP pair (the generic kernels take MO orbitals
// per species: 4 for s/p species, 9 with d shells)
template <int MO>
__global__ void k_svnl(DView v, DPP q) {
  constexpr int NB = MO * MO;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.npairs[PPL_PP]) return;
  const int ia = q.atom[PPL_PP][p];
  const int s = v.atom_struct[ia];
  const int ja = v.st[s].atom_off + q.j[PPL_PP][p] - 1;
  const int mb = q.b[PPL_PP][p];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int in1 = v.imass[ia], in2 = v.imass[ja];
  const int nmu = v.sp.norb[in1 - 1], nnu = q.norb_pp[in2 - 1];
  double r1[3], r2[3], sighat[3], eps[9], pm[9];
  for (int k = 0; k < 3; k++) {
    r1[k] = v.ratom[3 * ia + k];
    r2[k] = v.ratom[3 * ja + k] + xls[3 * mb + k];
  }
  const double z = d_distance(r1, r2);
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = (r2[k] - r1[k]) / z;
  d_epsilon(r2, sighat, eps);
  d_pmat<double>(eps, pm);
  double dm[MO == FB_MOD ? 25 : 1];
  if (MO == FB_MOD) d_dmat(eps, dm);
  const int cid = v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + SLOT_VNL) * FB_MAXSUB2];
  int status = 0;
  double hm[NB], X[NB];
  if (cid < 0) {
    status = ST_NO_TABLE;
  } else {
    const DCell2c c = v.cell2c[cid];
    double F[NB];
    status = (c.nME > NB) ? ST_CAPACITY : d_interp2c(c, v.tab2c, z, F);
    if (!status) {
      for (int i = 0; i < nmu * nnu; i++) hm[i] = 0.0;
      for (int k = 0; k < c.nME; k++) hm[(v.nu[c.munu + k] - 1) * nmu + (v.mu[c.munu + k] - 1)] = F[k];
      d_rotate_pp(v.sp, q, in1, in2, pm, dm, hm, nmu, X);
      double *dst = q.svnl + q.sv_off[p];
      for (int i = 0; i < nmu * nnu; i++) dst[i] = X[i];
    }
  }
  if (status) atomicOr(v.status, status);
}

// X(mu, nu) = sum_c cl(c) * A(mu, c) * B(nu, c), c ascending from 0 (the
// PPx / bcnlx loops of M_assemble_PP_2c.f90 and M_assemble_PP_3c.f90);
// A is nmu x ncc, B nnu x ncc, column-major
__device__ __forceinline__ void kb_contract(int nmu, int nnu, int in_pp, const DPP &q, const double *A,
                                            const double *B, double *X) {
  double cl[NBP];
  int ncc = 0;
  for (int sh = 0; sh < q.nssh_pp[in_pp - 1]; sh++)
    for (int m = 0; m < 2 * q.lssh_pp[in_pp - 1][sh] + 1; m++) cl[ncc++] = q.cl[in_pp - 1][sh];
  for (int inu = 0; inu < nnu; inu++)
    for (int imu = 0; imu < nmu; imu++) {
      double x = 0.0;
      for (int c = 0; c < ncc; c++) x = x + cl[c] * A[c * nmu + imu] * B[c * nnu + inu];
      X[inu * nmu + imu] = x;
    }
}

// assemble_vnl_2c, atom case: the self block of iatom sums, over its PP
// neighbours in order, sum_c cl(c) sVNL(mu, c) sVNL(nu, c) of that neighbour's
// projectors; one lane per (atom, block element), each element summed in the
// reference's order.
template <int MO>
__global__ void k_vnl_2c_self(DView v, DPP q) {
  constexpr int NB = MO * MO;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)v.natoms * NB) return;
  const int ia = (int)(t / NB), e = (int)(t % NB);
  const int in1 = v.imass[ia];
  const int nmu = v.sp.norb[in1 - 1];
  if (e >= nmu * nmu) return;
  const int imu = e % nmu, inu = e / nmu;
  const int aoff = v.st[v.atom_struct[ia]].atom_off;
  double acc = 0.0;
  for (int64_t p = q.off[PPL_PP][ia]; p < q.off[PPL_PP][ia + 1]; p++) {
    const int inp = v.imass[aoff + q.j[PPL_PP][p] - 1];
    const double *sv = q.svnl + q.sv_off[p];
    double x = 0.0;
    int c = 0;
    for (int sh = 0; sh < q.nssh_pp[inp - 1]; sh++) {
      const double cl = q.cl[inp - 1][sh];
      for (int m = 0; m < 2 * q.lssh_pp[inp - 1][sh] + 1; m++, c++) x = x + cl * sv[c * nmu + imu] * sv[c * nmu + inu];
    }
    acc = acc + x;
  }
  q.vnl[q.vl_off[q.off[PPL_PPP][ia] + q.self[PPL_PPP][ia] - 1] + e] = acc;
}

// assemble_vnl_2c, ontop cases: one lane per non-self PPp pair, its ontop-left
// term (the PPx pair mapped onto it, projectors of iatom) and then its
// ontop-right term (the PP pair mapped onto it, projectors of jatom), as the
// reference's two loops add them.
template <int MO>
__global__ void k_vnl_2c(DView v, DPP q) {
  constexpr int NB = MO * MO;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= q.npairs[PPL_PPP]) return;
  const int ia = q.atom[PPL_PPP][k];
  if (k == q.off[PPL_PPP][ia] + q.self[PPL_PPP][ia] - 1) return;
  const int s = v.atom_struct[ia];
  const int aoff = v.st[s].atom_off;
  const int ja = aoff + q.j[PPL_PPP][k] - 1;
  const int in1 = v.imass[ia], in2 = v.imass[ja];
  const int nmu = v.sp.norb[in1 - 1], nnu = v.sp.norb[in2 - 1], nb = nmu * nnu;
  double acc[NB], X[NB];
  for (int i = 0; i < nb; i++) acc[i] = 0.0;
  const int64_t o0 = q.off[PPL_PP][ia];
  const int px = q.inv_px[k];
  if (px >= 0) {
    const double *sv1 = q.svnl + q.sv_off[o0 + q.self[PPL_PP][ia] - 1];
    const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][ja] + q.px_point[px] - 1];
    kb_contract(nmu, nnu, in1, q, sv1, sv2, X);
    for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i];
  }
  const int pp = q.inv_pp[k];
  if (pp >= 0) {
    const double *sv1 = q.svnl + q.sv_off[pp];
    const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][ja] + q.self[PPL_PP][ja] - 1];
    kb_contract(nmu, nnu, in2, q, sv1, sv2, X);
    for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i];
  }
  double *dst = q.vnl + q.vl_off[k];
  for (int i = 0; i < nb; i++) dst[i] = acc[i];
}

// mpairnay (M_assemble_PP_3c.f90:266-364): the one PP neighbour of iatom that is
// jatom (1-based local) at offset r12 from iatom's own position; 0 if not unique
__device__ inline int mpairnay(const DView &v, const DPP &q, const double *xls, int aoff, int gia, int lja,
                               const double *r12) {
  int imatch = 0, mpair = 0;
  const int64_t o = q.off[PPL_PP][gia];
  const int gja = aoff + lja - 1;
  for (int k = 0; k < q.n[PPL_PP][gia]; k++) {
    if (q.j[PPL_PP][o + k] != lja) continue;
    const int mb = q.b[PPL_PP][o + k];
    double r21[3];
#pragma unroll
    for (int c = 0; c < 3; c++) r21[c] = (v.ratom[3 * gja + c] + xls[3 * mb + c]) - v.ratom[3 * gia + c];
    if (d_distance(r12, r21) < 0.0001) {
      imatch++;
      mpair = k + 1;
    }
  }
  return imatch == 1 ? mpair : 0;
}

// the two mpairnay searches of every sorted triplet, done once per list build
// (they depend on the lists and positions only)
__global__ void k_pp_mpair(DView v, DPP q) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= q.ntrip) return;
  const int4 tr = q.trip_s[t];
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const int aoff = v.st[s].atom_off;
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r31[3], r32[3];
  for (int k = 0; k < 3; k++) {
    const double rna = v.ratom[3 * ial + k];
    r31[k] = rna - (v.ratom[3 * gia + k] + xls[3 * ib + k]);
    r32[k] = rna - (v.ratom[3 * gja + k] + xls[3 * jb + k]);
  }
  const int lal = ial - aoff + 1;
  q.mp[t] = make_int2(mpairnay(v, q, xls, aoff, gia, lal, r31), mpairnay(v, q, xls, aoff, gja, lal, r32));
}

// assemble_vnl_3c: the workgroup owns the target segments that start in its
// tile of sorted triplets; each thread forms bcnlx of one triplet per chunk
// into LDS, one owner per segment adds them in list order onto the block
template <int MO>
__global__ __launch_bounds__(T3P) void k_vnl_3c(DView v, DPP q) {
  constexpr int NB = MO * MO;
  __shared__ double s_c[T3P][NB + 1];
  const int64_t ts = xcd_tile(blockIdx.x, gridDim.x) * T3P;
  const int64_t te = min(ts + T3P, q.ntrip);
  const int64_t sgs = q.trip_seg[ts];
  const int64_t first = (q.seg_off[sgs] == ts) ? sgs : sgs + 1;
  const int64_t last = q.trip_seg[te - 1];
  if (first > last) return;
  const int64_t r0 = q.seg_off[first], r1 = q.seg_off[last + 1];
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  int64_t a = 0, b = 0, dst = 0;
  int nb = 0;
  double acc[NB];
  if (own) {
    a = q.seg_off[my];
    b = q.seg_off[my + 1];
    const int k = q.seg_pair[my];
    const int ia = q.atom[PPL_PPP][k];
    const int ja = v.st[v.atom_struct[ia]].atom_off + q.j[PPL_PPP][k] - 1;
    nb = v.sp.norb[v.imass[ia] - 1] * v.sp.norb[v.imass[ja] - 1];
    dst = q.vl_off[k];
    for (int i = 0; i < NB; i++) acc[i] = (i < nb) ? q.vnl[dst + i] : 0.0;
  }
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3P) {
    const int64_t t = c0 + threadIdx.x;
    double X[NB];
    for (int i = 0; i < NB; i++) X[i] = 0.0;
    if (t < r1) {
      const int4 tr = q.trip_s[t];
      const int ial = tr.x, gia = tr.y, gja = tr.z;
      const int2 mp = q.mp[t];
      const int m31 = mp.x, m32 = mp.y;
      if (m31 == 0 || m32 == 0) {
        status |= ST_BACK;
      } else {
        const int nmu = v.sp.norb[v.imass[gia] - 1], nnu = v.sp.norb[v.imass[gja] - 1];
        const double *sv1 = q.svnl + q.sv_off[q.off[PPL_PP][gia] + m31 - 1];
        const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][gja] + m32 - 1];
        kb_contract(nmu, nnu, v.imass[ial], q, sv1, sv2, X);
      }
    }
    for (int i = 0; i < NB; i++) s_c[threadIdx.x][i] = X[i];
    __syncthreads();
    if (own) {
      const int64_t k0 = max(a, c0), k1 = min(b, c0 + T3P);
      for (int64_t kk = k0; kk < k1; kk++)
        for (int i = 0; i < NB; i++)
          if (i < nb) acc[i] = acc[i] + s_c[kk - c0][i];
    }
    __syncthreads();
  }
  if (own)
    for (int i = 0; i < NB; i++)
      if (i < nb) q.vnl[dst + i] = acc[i];
  if (status) atomicOr(v.status, status);
}

// ---------------------------------------------------------------------------
// s/p species and s/p PP shells with canonical P_vnl tables: the same sums
// with compile-time block shapes.  The projector table of an (orbital, PP)
// pair has the element order of the 2c s/p tables (make_munu with the PP
// shells on the right) and rotate_PP is rotate, so the s/p interpolation and
// rotation of the 2c kernels apply unchanged; the contractions are unrolled.
#include "fb_sp.h"
#include "fb_2c_dev.h"

template <int CLS>
__device__ __forceinline__ int svnl_sp(const DView &v, int cid, double z, const double *pm, double *X) {
  if (cid < 0) return ST_NO_TABLE;
  const DCell2c c = v.cell2c[cid];
  if (z < 0.0) return ST_X_NEG;
  if (z > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  const DGrid2c gr = v.grid2c[c.grid];
  const int ileft = window7(z, gr.dx, gr.nx);
  double prod[7], F[nme2c_of(CLS)];
  lagrange7(v, gr, z, ileft, prod);
  interp_cell<nme2c_of(CLS)>(v, c, ileft, prod, F);
  rot2c_sp<double>(CLS, pm, F, X);
  return 0;
}

__global__ void k_svnl_sp(DView v, DPP q) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= q.npairs[PPL_PP]) return;
  const int ia = q.atom[PPL_PP][p];
  const int s = v.atom_struct[ia];
  const int ja = v.st[s].atom_off + q.j[PPL_PP][p] - 1;
  const int mb = q.b[PPL_PP][p];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int in1 = v.imass[ia], in2 = v.imass[ja];
  double r1[3], r2[3], sighat[3], eps[9], pm[9];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r1[k] = v.ratom[3 * ia + k];
    r2[k] = v.ratom[3 * ja + k] + xls[3 * mb + k];
  }
  const double z = d_distance(r1, r2);
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else {
#pragma unroll
    for (int k = 0; k < 3; k++) sighat[k] = (r2[k] - r1[k]) / z;
  }
  d_epsilon(r2, sighat, eps);
  d_pmat<double>(eps, pm);
  const int cid = v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + SLOT_VNL) * FB_MAXSUB2];
  const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (q.norb_pp[in2 - 1] == 4 ? 1 : 0);
  double X[16];
  int status;
  double *dst = q.svnl + q.sv_off[p];
  switch (cls) {
    case 3:
      status = svnl_sp<3>(v, cid, z, pm, X);
#pragma unroll
      for (int i = 0; i < 16; i++) dst[i] = X[i];
      break;
    case 2:
      status = svnl_sp<2>(v, cid, z, pm, X);
#pragma unroll
      for (int i = 0; i < 4; i++) dst[i] = X[i];
      break;
    case 1:
      status = svnl_sp<1>(v, cid, z, pm, X);
#pragma unroll
      for (int i = 0; i < 4; i++) dst[i] = X[i];
      break;
    default:
      status = svnl_sp<0>(v, cid, z, pm, X);
      dst[0] = X[0];
      break;
  }
  if (status) atomicOr(v.status, status);
}

// X(mu, nu) = sum_c cl(c) A(mu, c) B(nu, c), unrolled, the order of kb_contract
template <int NMU, int NNU, int NCC>
__device__ __forceinline__ void kb_sp(const double *cl, const double *A, const double *B, double *X) {
#pragma unroll
  for (int inu = 0; inu < NNU; inu++)
#pragma unroll
    for (int imu = 0; imu < NMU; imu++) {
      double x = 0.0;
#pragma unroll
      for (int c = 0; c < NCC; c++) x = x + cl[c] * A[c * NMU + imu] * B[c * NNU + inu];
      X[inu * NMU + imu] = x;
    }
}

template <int NMU, int NNU, int NCC>
__device__ __forceinline__ void kb_add(const DPP &q, int in_pp, const double *A, const double *B, double *acc) {
  double cl[NCC], X[NMU * NNU];
#pragma unroll
  for (int c = 0; c < NCC; c++) cl[c] = q.clo[in_pp - 1][c];
  kb_sp<NMU, NNU, NCC>(cl, A, B, X);
#pragma unroll
  for (int i = 0; i < NMU * NNU; i++) acc[i] = acc[i] + X[i];
}

template <int NMU, int NNU>
__device__ __forceinline__ void kb_add_any(const DPP &q, int in_pp, const double *A, const double *B, double *acc) {
  if (q.norb_pp[in_pp - 1] == 4) kb_add<NMU, NNU, 4>(q, in_pp, A, B, acc);
  else kb_add<NMU, NNU, 1>(q, in_pp, A, B, acc);
}

template <int NMU, int NNU>
__device__ __forceinline__ void vnl_pair_sp(const DView &v, const DPP &q, int64_t k, int ia, int ja, int in1, int in2) {
  double acc[NMU * NNU];
#pragma unroll
  for (int i = 0; i < NMU * NNU; i++) acc[i] = 0.0;
  const int64_t o0 = q.off[PPL_PP][ia];
  const int px = q.inv_px[k];
  if (px >= 0) {
    const double *sv1 = q.svnl + q.sv_off[o0 + q.self[PPL_PP][ia] - 1];
    const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][ja] + q.px_point[px] - 1];
    kb_add_any<NMU, NNU>(q, in1, sv1, sv2, acc);
  }
  const int pp = q.inv_pp[k];
  if (pp >= 0) {
    const double *sv1 = q.svnl + q.sv_off[pp];
    const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][ja] + q.self[PPL_PP][ja] - 1];
    kb_add_any<NMU, NNU>(q, in2, sv1, sv2, acc);
  }
  double *dst = q.vnl + q.vl_off[k];
#pragma unroll
  for (int i = 0; i < NMU * NNU; i++) dst[i] = acc[i];
}

// atom case of assemble_vnl_2c for s/p PP shells: k_vnl_2c_self's sums (same
// terms, same order) with the loads of four PP neighbours in flight per step.
// The projector count of a block comes from its sVNL offsets, so the block
// loads wait for sv_off only, beside the neighbour-species chain.
__global__ void k_vnl_2c_self_sp(DView v, DPP q) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)v.natoms * NBP) return;
  const int ia = (int)(t / NBP), e = (int)(t % NBP);
  const int nmu = v.sp.norb[v.imass[ia] - 1];
  if (e >= nmu * nmu) return;
  const int imu = e % nmu, inu = e / nmu;
  const int aoff = v.st[v.atom_struct[ia]].atom_off;
  const int64_t p0 = q.off[PPL_PP][ia], p1 = q.off[PPL_PP][ia + 1];
  double acc = 0.0;
  for (int64_t pb = p0; pb < p1; pb += 4) {
    int inp[4], npp[4];
    double A[4][4], B[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t p = pb + u;
      inp[u] = 1;
      npp[u] = 0;
      if (p < p1) {
        const int64_t so = q.sv_off[p];
        npp[u] = (int)(q.sv_off[p + 1] - so) / nmu;
        inp[u] = v.imass[aoff + q.j[PPL_PP][p] - 1];
        const double *sv = q.svnl + so;
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (c < npp[u]) {
            A[u][c] = sv[c * nmu + imu];
            B[u][c] = sv[c * nmu + inu];
          }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (pb + u < p1) {
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (c < npp[u]) x = x + q.clo[inp[u] - 1][c] * A[u][c] * B[u][c];
        acc = acc + x;
      }
  }
  q.vnl[q.vl_off[q.off[PPL_PPP][ia] + q.self[PPL_PPP][ia] - 1] + e] = acc;
}

__global__ void k_vnl_2c_sp(DView v, DPP q) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= q.npairs[PPL_PPP]) return;
  const int ia = q.atom[PPL_PPP][k];
  if (k == q.off[PPL_PPP][ia] + q.self[PPL_PPP][ia] - 1) return;
  const int ja = v.st[v.atom_struct[ia]].atom_off + q.j[PPL_PPP][k] - 1;
  const int in1 = v.imass[ia], in2 = v.imass[ja];
  const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
  switch (cls) {
    case 3: vnl_pair_sp<4, 4>(v, q, k, ia, ja, in1, in2); break;
    case 2: vnl_pair_sp<4, 1>(v, q, k, ia, ja, in1, in2); break;
    case 1: vnl_pair_sp<1, 4>(v, q, k, ia, ja, in1, in2); break;
    default: vnl_pair_sp<1, 1>(v, q, k, ia, ja, in1, in2); break;
  }
}

// bcnlx of one triplet into its LDS row
template <int NMU, int NNU>
__device__ __forceinline__ void triplet_vnl_sp(const DPP &q, int indna, const double *sv1, const double *sv2,
                                               double *row) {
  double X[NMU * NNU];
  if (q.norb_pp[indna - 1] == 4) {
    double cl[4];
#pragma unroll
    for (int c = 0; c < 4; c++) cl[c] = q.clo[indna - 1][c];
    kb_sp<NMU, NNU, 4>(cl, sv1, sv2, X);
  } else {
    const double cl[1] = {q.clo[indna - 1][0]};
    kb_sp<NMU, NNU, 1>(cl, sv1, sv2, X);
  }
#pragma unroll
  for (int i = 0; i < NMU * NNU; i++) row[i] = X[i];
}

// The workgroup owns the whole segments of one tile from the list build (at
// most T3P triplets unless one segment is longer), so it usually makes a
// single pass over its triplets.
__global__ __launch_bounds__(T3P) void k_vnl_3c_sp(DView v, DPP q) {
  __shared__ double s_c[T3P][NBP + 1];
  const int4 td = q.tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int64_t first = td.x, last = td.y, r0 = td.z, r1 = td.w;
  if (first > last) return;
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  int64_t a = 0, b = 0, dst = 0;
  int nb = 0;
  double acc[NBP];
#pragma unroll
  for (int i = 0; i < NBP; i++) acc[i] = 0.0;
  if (own) {
    a = q.seg_off[my];
    b = q.seg_off[my + 1];
    const int k = q.seg_pair[my];
    const int ia = q.atom[PPL_PPP][k];
    const int ja = v.st[v.atom_struct[ia]].atom_off + q.j[PPL_PPP][k] - 1;
    nb = v.sp.norb[v.imass[ia] - 1] * v.sp.norb[v.imass[ja] - 1];
    dst = q.vl_off[k];
#pragma unroll
    for (int i = 0; i < NBP; i++)
      if (i < nb) acc[i] = q.vnl[dst + i];
  }
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3P) {
    const int64_t t = c0 + threadIdx.x;
    double *row = s_c[threadIdx.x];
    if (t < r1) {
      const int4 tr = q.trip_s[t];
      const int ial = tr.x, gia = tr.y, gja = tr.z;
      const int2 mp = q.mp[t];
      if (mp.x == 0 || mp.y == 0) {
        status |= ST_BACK;
      } else {
        const int in1 = v.imass[gia], in2 = v.imass[gja];
        const double *sv1 = q.svnl + q.sv_off[q.off[PPL_PP][gia] + mp.x - 1];
        const double *sv2 = q.svnl + q.sv_off[q.off[PPL_PP][gja] + mp.y - 1];
        const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
        switch (cls) {
          case 3: triplet_vnl_sp<4, 4>(q, v.imass[ial], sv1, sv2, row); break;
          case 2: triplet_vnl_sp<4, 1>(q, v.imass[ial], sv1, sv2, row); break;
          case 1: triplet_vnl_sp<1, 4>(q, v.imass[ial], sv1, sv2, row); break;
          default: triplet_vnl_sp<1, 1>(q, v.imass[ial], sv1, sv2, row); break;
        }
      }
    }
    __syncthreads();
    if (own) {
      const int64_t k0 = max(a, c0), k1 = min(b, c0 + T3P);
      for (int64_t kk = k0; kk < k1; kk++)
#pragma unroll
        for (int i = 0; i < NBP; i++)
          if (i < nb) acc[i] = acc[i] + s_c[kk - c0][i];
    }
    __syncthreads();
  }
  if (own) {
#pragma unroll
    for (int i = 0; i < NBP; i++)
      if (i < nb) q.vnl[dst + i] = acc[i];
  }
  if (status) atomicOr(v.status, status);
}

// ------------------------------------------------------------------ launchers
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" {
int fbk_pp_neighbors(const DView &v, const DPP &q, int list, int fill, int *status, hipStream_t st) {
  (void)status;
  if (v.natoms == 0) return 0;
  const size_t lds = sizeof(unsigned) * (size_t)v.cell_bits;
  if (fill)
    hipLaunchKernelGGL(k_pp_neighbors<true>, dim3(v.natoms), dim3(LWG), lds, st, v, q, list);
  else
    hipLaunchKernelGGL(k_pp_neighbors<false>, dim3(v.natoms), dim3(LWG), lds, st, v, q, list);
  return (int)hipGetLastError();
}

int fbk_ppp_candidates(const DView &v, const DPP &q, int64_t *cnt, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_ppp_candidates, dim3(nblk(v.natoms, 128)), dim3(128), 0, st, v, q, cnt);
  return (int)hipGetLastError();
}

int fbk_ppp_build(const DView &v, const DPP &q, int fill, int *status, hipStream_t st) {
  if (v.natoms == 0) return 0;
  if (fill) {
    hipLaunchKernelGGL(k_ppp_build<true>, dim3(v.natoms), dim3(LWG), 0, st, v, q, status);
  } else {
    hipLaunchKernelGGL(k_ppp_keys, dim3(v.natoms), dim3(LWG), 0, st, v, q);
    hipLaunchKernelGGL(k_ppp_build<false>, dim3(v.natoms), dim3(LWG), 0, st, v, q, status);
  }
  return (int)hipGetLastError();
}

int fbk_pp_maps(const DView &v, const DPP &q, int *status, hipStream_t st) {
  const int64_t n = q.npairs[PPL_PP] + q.npairs[PPL_PPX];
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_pp_maps, dim3(nblk(n, 128)), dim3(128), 0, st, v, q, status);
  return (int)hipGetLastError();
}

int fbk_ppk_size(const DView &v, const DPP &q, int *sz, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_ppk_size, dim3(nblk(v.natoms, 256)), dim3(256), 0, st, v, q, sz);
  return (int)hipGetLastError();
}

int fbk_ppk_insert(const DView &v, const DPP &q, hipStream_t st) {
  if (q.npairs[PPL_PPP] == 0) return 0;
  hipLaunchKernelGGL(k_ppk_insert, dim3(nblk(q.npairs[PPL_PPP], 256)), dim3(256), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_ppp_disp(const DView &v, const DPP &q, hipStream_t st) {
  if (q.npairs[PPL_PPP] == 0) return 0;
  hipLaunchKernelGGL(k_ppp_disp, dim3(nblk(q.npairs[PPL_PPP], 256)), dim3(256), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_pp_common(const DView &v, const DPP &q, int fill, hipStream_t st) {
  if (v.natoms == 0) return 0;
  if (fill)
    hipLaunchKernelGGL(k_pp_common<true>, dim3(v.natoms), dim3(LWG), 0, st, v, q);
  else
    hipLaunchKernelGGL(k_pp_common<false>, dim3(v.natoms), dim3(LWG), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_pp_block_sizes(const DView &v, const DPP &q, int which, int64_t *sz, hipStream_t st) {
  const int64_t n = q.npairs[which == 0 ? PPL_PP : PPL_PPP];
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_pp_block_sizes, dim3(nblk(n, 256)), dim3(256), 0, st, v, q, which, sz);
  return (int)hipGetLastError();
}

int fbk_svnl(const DView &v, const DPP &q, hipStream_t st) {
  if (q.npairs[PPL_PP] == 0) return 0;
  if (q.canon)
    hipLaunchKernelGGL(k_svnl_sp, dim3(nblk(q.npairs[PPL_PP], 128)), dim3(128), 0, st, v, q);
  else if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_svnl<FB_MOD>, dim3(nblk(q.npairs[PPL_PP], 128)), dim3(128), 0, st, v, q);
  else
    hipLaunchKernelGGL(k_svnl<FB_MO>, dim3(nblk(q.npairs[PPL_PP], 128)), dim3(128), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_vnl_2c(const DView &v, const DPP &q, hipStream_t st) {
  if (q.npairs[PPL_PPP] == 0) return 0;
  if (q.canon)
    hipLaunchKernelGGL(k_vnl_2c_self_sp, dim3(nblk((int64_t)v.natoms * NBP, 256)), dim3(256), 0, st, v, q);
  else if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_vnl_2c_self<FB_MOD>, dim3(nblk((int64_t)v.natoms * FB_MOD * FB_MOD, 256)), dim3(256), 0, st,
                       v, q);
  else
    hipLaunchKernelGGL(k_vnl_2c_self<FB_MO>, dim3(nblk((int64_t)v.natoms * NBP, 256)), dim3(256), 0, st, v, q);
  if (q.canon)
    hipLaunchKernelGGL(k_vnl_2c_sp, dim3(nblk(q.npairs[PPL_PPP], 128)), dim3(128), 0, st, v, q);
  else if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_vnl_2c<FB_MOD>, dim3(nblk(q.npairs[PPL_PPP], 128)), dim3(128), 0, st, v, q);
  else
    hipLaunchKernelGGL(k_vnl_2c<FB_MO>, dim3(nblk(q.npairs[PPL_PPP], 128)), dim3(128), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_pp_mpair(const DView &v, const DPP &q, hipStream_t st) {
  if (q.ntrip == 0) return 0;
  hipLaunchKernelGGL(k_pp_mpair, dim3(nblk(q.ntrip, 128)), dim3(128), 0, st, v, q);
  return (int)hipGetLastError();
}

int fbk_vnl_3c(const DView &v, const DPP &q, hipStream_t st) {
  if (q.ntrip == 0) return 0;
  if (q.canon) {
    if (q.ntiles > 0) hipLaunchKernelGGL(k_vnl_3c_sp, dim3((unsigned)q.ntiles), dim3(T3P), 0, st, v, q);
  } else if (fb_wide_species(v.sp)) {
    hipLaunchKernelGGL(k_vnl_3c<FB_MOD>, dim3(nblk(q.ntrip, T3P)), dim3(T3P), 0, st, v, q);
  } else {
    hipLaunchKernelGGL(k_vnl_3c<FB_MO>, dim3(nblk(q.ntrip, T3P)), dim3(T3P), 0, st, v, q);
  }
  return (int)hipGetLastError();
}
}

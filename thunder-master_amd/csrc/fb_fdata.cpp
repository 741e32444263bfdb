// Reader of the on-disk Fdata tables (the text layout that read_Fdata_2c,
// e.FDATA/M_Fdata_2c.f90:149-272, and read_Fdata_3c, e.FDATA/M_Fdata_3c.f90:143-264,
// consume), feeding fb_set_table_2c / fb_set_table_3c.
//
// Files per ordered species pair (Z1, Z2) and triple (Z1, Z2, Z3), names with
// two-digit atomic numbers:
//   2c.Z1.Z2.dat         number of cells
//   2c.Z1.Z2.dir         per cell: iindex itype isubtype filename nME nx xmax
//   <filename>           nx records of nME values (grid point major)
//   TT_munu_2c.Z1.Z2.dat mu, nu, mvalue records of the cells of interaction TT
//   3c.Z1.Z2.Z3.dat/.dir per cell: iindex itype isubtype itheta filename nME nx xmax ny ymax
//   <filename>           ny*nx records (iy outer, ix inner) of nME values
//   TT_munu_3c.Z1.Z2.Z3.dat
// Input is list-directed: each read starts at a new record and takes values
// across records until it has them all; separators are blanks and commas,
// exponents may be written with d/D, and r*c repeats c r times.
//
// Reference behaviours kept on purpose:
//  * a cell's (itype, isubtype[, itheta]) maps to the cell stored at position
//    iindex of the directory, not at the record's own position;
//  * the mu/nu/mvalue unit stays open across cells: consecutive cells of one
//    interaction take consecutive triples of records; switching to another
//    interaction's file closes the previous one, and a later return to it
//    starts from its first record (as the Fortran standard defines OPEN on a
//    connected unit; directories written by create.x keep each interaction's
//    cells together, so this case does not arise there).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "fb_fdata.h"
#include "../../include/fireball_hip.h"

namespace {

// list-directed record reader over one text file
class ListReader {
 public:
  bool open(const std::string &path) {
    in_.close();
    in_.clear();
    in_.open(path);
    path_ = path;
    return (bool)in_;
  }
  const std::string &path() const { return path_; }
  bool is_open() const { return in_.is_open(); }

  // the next n values, starting at a new record
  bool values(int n, std::vector<std::string> &out) {
    out.clear();
    std::string line;
    while ((int)out.size() < n) {
      if (!std::getline(in_, line)) return false;
      if (!split(line, n, out)) break;   // a '/' ends the read
    }
    return (int)out.size() >= n;
  }

 private:
  // appends the tokens of one record; false when a '/' terminates the read
  static bool split(const std::string &line, int n, std::vector<std::string> &out) {
    size_t i = 0;
    while (i < line.size() && (int)out.size() < n) {
      while (i < line.size() && (std::isspace((unsigned char)line[i]) || line[i] == ',')) i++;
      if (i >= line.size()) break;
      if (line[i] == '/') return false;
      std::string tok;
      if (line[i] == '\'' || line[i] == '"') {
        const char q = line[i++];
        while (i < line.size() && line[i] != q) tok += line[i++];
        if (i < line.size()) i++;
      } else {
        while (i < line.size() && !std::isspace((unsigned char)line[i]) && line[i] != ',' && line[i] != '/')
          tok += line[i++];
      }
      const size_t star = tok.find('*');
      if (star != std::string::npos && star > 0 && std::isdigit((unsigned char)tok[0])) {
        const int r = std::atoi(tok.substr(0, star).c_str());
        const std::string v = tok.substr(star + 1);
        for (int k = 0; k < r && (int)out.size() < n; k++) out.push_back(v);
      } else {
        out.push_back(tok);
      }
    }
    return true;
  }

  std::ifstream in_;
  std::string path_;
};

bool to_int(const std::string &s, int &v) {
  char *end = nullptr;
  long x = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end) return false;
  v = (int)x;
  return true;
}

bool to_double(std::string s, double &v) {
  for (char &c : s)
    if (c == 'd' || c == 'D') c = 'e';
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end != s.c_str() && !*end;
}

std::string zname(const char *fmt, int a, int b, int c = 0, int d = 0) {
  char buf[96];
  std::snprintf(buf, sizeof buf, fmt, a, b, c, d);
  return buf;
}

struct CellMeta {
  int iindex, itype, isub, itheta, nME, nx, ny;
  double xmax, ymax;
  std::string file;
  std::vector<int> mu, nu, mv;
  std::vector<double> data;
};

bool read_munu(ListReader &u, const std::string &path, CellMeta &c, std::string &err) {
  if (!u.is_open() || u.path() != path) {
    if (!u.open(path)) {
      err = "cannot open " + path;
      return false;
    }
  }
  std::vector<std::string> t;
  std::vector<int> *dst[3] = {&c.mu, &c.nu, &c.mv};
  for (auto *d : dst) {
    if (!u.values(c.nME, t)) {
      err = "short record in " + path;
      return false;
    }
    d->resize(c.nME);
    for (int k = 0; k < c.nME; k++)
      if (!to_int(t[k], (*d)[k])) {
        err = "bad integer in " + path;
        return false;
      }
  }
  return true;
}

bool read_data(const std::string &path, int nrec, int nME, std::vector<double> &out, std::string &err) {
  ListReader r;
  if (!r.open(path)) {
    err = "cannot open " + path;
    return false;
  }
  out.resize((size_t)nrec * nME);
  std::vector<std::string> t;
  for (int i = 0; i < nrec; i++) {
    if (!r.values(nME, t)) {
      err = "short data in " + path;
      return false;
    }
    for (int m = 0; m < nME; m++)
      if (!to_double(t[m], out[(size_t)i * nME + m])) {
        err = "bad number in " + path;
        return false;
      }
  }
  return true;
}

bool read_count(const std::string &path, int &n, std::string &err) {
  ListReader r;
  std::vector<std::string> t;
  if (!r.open(path) || !r.values(1, t) || !to_int(t[0], n)) {
    err = "cannot read the cell count of " + path;
    return false;
  }
  return true;
}

}  // namespace

int fb_fdata_read(const std::string &dir, int ns, const int *nZ, const FdataWanted &want, FdataSink &sink,
                  std::string &err) {
  for (int a = 1; a <= ns; a++)
    for (int b = 1; b <= ns; b++) {
      // ---- two-centre cells of the ordered pair (a, b)
      const int za = nZ[a - 1], zb = nZ[b - 1];
      int ncell = 0;
      if (!read_count(dir + zname("/2c.%02d.%02d.dat", za, zb), ncell, err)) return -1;
      ListReader d, u;
      const std::string dpath = dir + zname("/2c.%02d.%02d.dir", za, zb);
      if (!d.open(dpath)) {
        err = "cannot open " + dpath;
        return -1;
      }
      std::vector<CellMeta> cells(ncell);
      std::vector<std::string> t;
      for (int ic = 0; ic < ncell; ic++) {
        CellMeta &c = cells[ic];
        if (!d.values(7, t) || !to_int(t[0], c.iindex) || !to_int(t[1], c.itype) || !to_int(t[2], c.isub) ||
            !to_int(t[4], c.nME) || !to_int(t[5], c.nx) || !to_double(t[6], c.xmax) || c.nME < 1 || c.nx < 1) {
          err = "bad record in " + dpath;
          return -1;
        }
        c.file = t[3];
        c.itheta = 0;
        c.ny = 1;
        c.ymax = 0.0;
        if (want.type2c(c.itype)) {
          std::vector<double> raw;
          if (!read_data(dir + "/" + c.file, c.nx, c.nME, raw, err)) return -1;
          // records are grid points: [ix][ME] -> Fortran Fdata_2c(ix, ME)
          c.data.resize(raw.size());
          for (int ix = 0; ix < c.nx; ix++)
            for (int m = 0; m < c.nME; m++) c.data[(size_t)m * c.nx + ix] = raw[(size_t)ix * c.nME + m];
        }
        if (!read_munu(u, dir + zname("/%02d_munu_2c.%02d.%02d.dat", c.itype, za, zb), c, err)) return -1;
      }
      // (itype, isubtype) -> the cell at directory position iindex
      std::map<std::pair<int, int>, int> index;
      for (auto &c : cells) index[{c.itype, c.isub}] = c.iindex;
      for (auto &kv : index) {
        if (!want.type2c(kv.first.first)) continue;
        const int pos = kv.second;
        if (pos < 1 || pos > ncell) {
          err = "2c index outside the directory in " + dpath;
          return -1;
        }
        const CellMeta &c = cells[pos - 1];
        if (!want.type2c(c.itype)) {
          err = "2c index points at a cell of another interaction in " + dpath;
          return -1;
        }
        const int rc = sink.cell2c(a, b, kv.first.first, kv.first.second, c.nME, c.nx, c.xmax, c.mu.data(),
                                   c.nu.data(), c.mv.data(), c.data.data());
        if (rc) return rc;
      }
      // ---- three-centre cells of the ordered triples (a, b, k)
      for (int k = 1; k <= ns; k++) {
        const int zk = nZ[k - 1];
        int n3 = 0;
        if (!read_count(dir + zname("/3c.%02d.%02d.%02d.dat", za, zb, zk), n3, err)) return -1;
        ListReader d3, u3;
        const std::string d3path = dir + zname("/3c.%02d.%02d.%02d.dir", za, zb, zk);
        if (!d3.open(d3path)) {
          err = "cannot open " + d3path;
          return -1;
        }
        std::vector<CellMeta> c3(n3);
        for (int ic = 0; ic < n3; ic++) {
          CellMeta &c = c3[ic];
          if (!d3.values(10, t) || !to_int(t[0], c.iindex) || !to_int(t[1], c.itype) || !to_int(t[2], c.isub) ||
              !to_int(t[3], c.itheta) || !to_int(t[5], c.nME) || !to_int(t[6], c.nx) ||
              !to_double(t[7], c.xmax) || !to_int(t[8], c.ny) || !to_double(t[9], c.ymax) || c.nME < 1 ||
              c.nx < 1 || c.ny < 1) {
            err = "bad record in " + d3path;
            return -1;
          }
          c.file = t[4];
          if (want.type3c(c.itype)) {
            std::vector<double> raw;
            if (!read_data(dir + "/" + c.file, c.ny * c.nx, c.nME, raw, err)) return -1;
            // records are (iy, ix) with ix fastest: [iy][ix][ME] -> Fortran
            // Fdata_3c(ix, iy, ME), the layout fb_set_table_3c takes
            c.data.resize(raw.size());
            for (int iy = 0; iy < c.ny; iy++)
              for (int ix = 0; ix < c.nx; ix++)
                for (int m = 0; m < c.nME; m++)
                  c.data[((size_t)m * c.ny + iy) * c.nx + ix] = raw[((size_t)iy * c.nx + ix) * c.nME + m];
          }
          if (!read_munu(u3, dir + zname("/%02d_munu_3c.%02d.%02d.%02d.dat", c.itype, za, zb, zk), c, err))
            return -1;
        }
        std::map<std::tuple<int, int, int>, int> index3;
        for (auto &c : c3) index3[std::make_tuple(c.itype, c.isub, c.itheta)] = c.iindex;
        for (auto &kv : index3) {
          if (!want.type3c(std::get<0>(kv.first))) continue;
          const int pos = kv.second;
          if (pos < 1 || pos > n3) {
            err = "3c index outside the directory in " + d3path;
            return -1;
          }
          const CellMeta &c = c3[pos - 1];
          if (!want.type3c(c.itype)) {
            err = "3c index points at a cell of another interaction in " + d3path;
            return -1;
          }
          const int rc = sink.cell3c(a, b, k, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                                     c.nME, c.nx, c.xmax, c.ny, c.ymax, c.mu.data(), c.nu.data(), c.mv.data(),
                                     c.data.data());
          if (rc) return rc;
        }
      }
    }
  return 0;
}

// Kleinman-Bylander coefficients: clPP.ZZ.dat holds one value per PP shell
// (e.FDATA/M_Fdata_2c.f90:251-258)
int fb_fdata_read_cl(const std::string &dir, int ns, const int *nZ, const int *nssh_pp, double *cl,
                     std::string &err) {
  for (int a = 0; a < ns; a++) {
    if (nssh_pp[a] == 0) continue;
    ListReader r;
    const std::string path = dir + zname("/clPP.%02d.dat", nZ[a], 0);
    if (!r.open(path)) {
      err = "cannot open " + path;
      return -1;
    }
    std::vector<std::string> t;
    for (int k = 0; k < nssh_pp[a]; k++)
      if (!r.values(1, t) || !to_double(t[0], cl[a * FB_MAXSH + k])) {
        err = "short or bad record in " + path;
        return -1;
      }
  }
  return 0;
}

// ---------------------------------------------------------------- positions
// The in-scope steps of read_positions (c.SYSTEM/M_configuraciones.f90:604-797)
// on a structure file in the reference's layout: "natoms icluster", three
// lattice vectors, then one "Z x y z" record per atom, all list-directed.
int fb_positions_read(const std::string &path, int ns, const int *nZ, const int *norb, int max_atoms, int *natoms,
                      int *icluster, double *lattice, double *ratom, int *imass, int *norbitals, int *iblock_slot,
                      int *shifted, std::string &err) {
  ListReader r;
  if (!r.open(path)) {
    err = "cannot open " + path;
    return -1;
  }
  std::vector<std::string> t;
  int n = 0, ic = 0;
  if (!r.values(2, t) || !to_int(t[0], n) || !to_int(t[1], ic) || n < 0) {
    err = "bad natoms/icluster record in " + path;
    return -1;
  }
  *natoms = n;
  *icluster = ic;
  if (n > max_atoms) {
    err = "more atoms than max_atoms in " + path;
    return -2;
  }
  for (int k = 0; k < 3; k++) {
    if (!r.values(3, t)) {
      err = "short lattice record in " + path;
      return -1;
    }
    for (int c = 0; c < 3; c++)
      if (!to_double(t[c], lattice[3 * k + c])) {
        err = "bad lattice value in " + path;
        return -1;
      }
  }
  for (int i = 0; i < n; i++) {
    int z = 0;
    if (!r.values(4, t) || !to_int(t[0], z)) {
      err = "bad atom record in " + path;
      return -1;
    }
    for (int c = 0; c < 3; c++)
      if (!to_double(t[1 + c], ratom[3 * i + c])) {
        err = "bad coordinate in " + path;
        return -1;
      }
    // species matching by atomic number: the first species with that Z (:666-681)
    int sp = 0;
    for (int k = 0; k < ns && !sp; k++)
      if (nZ[k] == z) sp = k + 1;
    if (!sp) {
      err = "atomic number " + std::to_string(z) + " of " + path + " is not among the species";
      return -3;
    }
    imass[i] = sp;
  }
  // norbitals (:703-708) and the orbital offset of each atom (:778-783)
  int no = 0;
  for (int i = 0; i < n; i++) {
    if (iblock_slot) iblock_slot[i] = no;
    no += norb[imass[i] - 1];
  }
  *norbitals = no;
  // shift every atom by (pi, 1/e, sqrt 2) when one lies within 1e-4 of the
  // origin (:757-775); magnitude is b.FUNCTIONS/magnitude.f90
  int shift = 0;
  for (int i = 0; i < n && !shift; i++) {
    const double *a = ratom + 3 * i;
    if (std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) < 1.0e-4) shift = 1;
  }
  if (shift) {
    const double shifter[3] = {4.0 * std::atan(1.0), 1.0 / std::exp(1.0), std::sqrt(2.0)};
    for (int i = 0; i < n; i++)
      for (int c = 0; c < 3; c++) ratom[3 * i + c] = ratom[3 * i + c] + shifter[c];
  }
  *shifted = shift;
  return 0;
}

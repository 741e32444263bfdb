// Host-side reader of the on-disk Fdata tables (see fb_fdata.cpp).
#pragma once

#include <string>

// which interactions to load (the others are skipped but their mu/nu/mvalue
// records are still consumed, as the reference's reader consumes them)
struct FdataWanted {
  bool (*type2c)(int itype);
  bool (*type3c)(int itype);
};

// receives every loaded cell, data in the Fortran layout of the reference's
// Fdata_2c(nx, nME) / Fdata_3c(nx, ny, nME); a nonzero return stops the read
struct FdataSink {
  virtual int cell2c(int is, int js, int itype, int isub, int nME, int nx, double xmax, const int *mu,
                     const int *nu, const int *mv, const double *fdata) = 0;
  virtual int cell3c(int is, int js, int ks, int itype, int isub, int itheta, int nME, int nx, double xmax, int ny,
                     double ymax, const int *mu, const int *nu, const int *mv, const double *fdata) = 0;
  virtual ~FdataSink() = default;
};

// reads every ordered species pair and triple of the ns species with atomic
// numbers nZ from directory dir; 0 on success, else -1 with err set (or the
// sink's return code)
int fb_fdata_read(const std::string &dir, int ns, const int *nZ, const FdataWanted &want, FdataSink &sink,
                  std::string &err);
int fb_fdata_read_cl(const std::string &dir, int ns, const int *nZ, const int *nssh_pp, double *cl,
                     std::string &err);

// the in-scope steps of read_positions on one structure file (see
// fb_fdata.cpp); 0, or -1 (format), -2 (capacity), -3 (unknown Z) with err set
int fb_positions_read(const std::string &path, int ns, const int *nZ, const int *norb, int max_atoms, int *natoms,
                      int *icluster, double *lattice, double *ratom, int *imass, int *norbitals, int *iblock_slot,
                      int *shifted, std::string &err);

// The spex.cor checkpoint file (host side, HDF5 C API): writes and reads
// the susceptibility, dielectric or inverse dielectric matrices of one
// k point in the layout of SPEX's writecor/readcor (correlation.f:
// 1636-1952), so that results are resumable and exchangeable with SPEX.
//
// Layout (COR version 1; the HDF5 wrapper conventions of Hwrapper.f):
// - file attributes "version" (int, 1), "nfreq,nfreqc" (int[2]), "freq"
//   (double[nfreq]), "freqc" (complex[nfreqc]);
// - one group per k point, named by the k vector in Fortran F12.10 format
//   joined by commas (correlation.f:1653, :1837);
// - group attributes "nbasm", "type", "label" (6 characters), "dim",
//   "ncblock", "oselfc", "contour" (double[4]) and, at the Gamma point,
//   "plasma";
// - group datasets "coulomb" (packed), "coul", "ctrafo", "cblock" (when
//   ncblock > 0), "matrix_<i>" (packed upper, i = 1..nfreq), "matrixc_<i>"
//   (nbas x nbas) and, at the Gamma point, "head", "wing", "headc",
//   "wingc".
// Every scalar attribute is a one-element array (hdf_rdwr_a of a scalar
// writes a rank-1 space of size 1).  A complex array is a double dataset
// with a leading dimension 2 (real, imaginary: Hwrapper.f hdf_rdwr_c), and
// Fortran dimensions appear reversed in HDF5's row-major view: a complex
// (nbas, nbas) matrix is the double dataset {nbas, nbas, 2}.
//
// Large datasets are divided as Hwrapper.f's hdf_rdwr divides them
// (Hwrapper.f:413-426, array_divide :573-585; cppmacro.h:143-144): an array
// of more than (2^31-1)/(element bytes) elements is written as "<name>_0"
// and "<name>_1", split at (ub-lb)/2+lb of its last Fortran dimension with
// more than one element, each half divided again by the same rule; reads
// follow the same rule.  The halves of a column-major array split along its
// slowest non-trivial dimension are contiguous.
//
// Built with -DINV, SPEX stores its MCOMPLEX_dp arrays (coulomb, ctrafo,
// matrix, head, wing) as real_dp; spexcor_set_inv selects that layout for
// writes and reads, and spexcor_file_inv detects it from the stored
// Coulomb matrix.
//
// COR version 0 files (no version attribute, or 0) are read as readcor reads
// them: one "matrix" dataset (all nfreq packed matrices) and one "matrixc"
// dataset (correlation.f:1799-1807).  The spex.cor.map directory of MPI
// runs (correlation.f:1655-1672, 1911-1946) maps a k point to the file
// spex.cor.<n> that holds it.
#include "../../include/spexcor.h"

#include <hdf5.h>

#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;
// MaxChunk of the HDF5 wrapper in bytes (huge(0)); settable for tests
size_t g_max_chunk = 2147483647;
// Layout of SPEX's MCOMPLEX_dp arrays (coulomb, ctrafo, matrix, head, wing):
// complex by default, real_dp in an inversion-symmetric (-DINV) build
// (cppmacro.h:50-60; correlation.f:89-96; global.f:153).
bool g_inv = false;

// dims of an MCOMPLEX_dp array: a trailing real/imaginary dimension unless
// the layout is the INV build's real one
std::vector<hsize_t> mc(std::vector<hsize_t> d) {
  if (!g_inv) d.push_back(2);
  return d;
}

int fail(const std::string &m) {
  g_err = m;
  return 1;
}

// Fortran F12.10 of one k coordinate (width 12, 10 decimals: a leading
// zero is dropped when the sign needs its place; overflow prints '*').
std::string f12_10(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.10f", v);
  std::string s(buf);
  if (s.size() > 12 && (s.rfind("-0.", 0) == 0)) s = "-" + s.substr(2);
  if (s.size() > 12) s = std::string(12, '*');
  if (s.size() < 12) s = std::string(12 - s.size(), ' ') + s;
  return s;
}

std::string group_name(const double k[3]) {
  return f12_10(k[0]) + "," + f12_10(k[1]) + "," + f12_10(k[2]);
}

struct H {  // closes an HDF5 handle on scope exit
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  operator hid_t() const { return id; }
};

// attributes ------------------------------------------------------------
bool write_attr(hid_t loc, const char *name, hid_t type, const void *buf, hsize_t n) {
  H spc(H5Screate_simple(1, &n, nullptr), H5Sclose);
  H att(H5Acreate2(loc, name, type, spc, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return spc >= 0 && att >= 0 && H5Awrite(att, type, buf) >= 0;
}

bool write_attr_cplx(hid_t loc, const char *name, const double *buf, hsize_t n) {
  hsize_t d[2] = {n, 2};
  H spc(H5Screate_simple(2, d, nullptr), H5Sclose);
  H att(H5Acreate2(loc, name, H5T_NATIVE_DOUBLE, spc, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return spc >= 0 && att >= 0 && H5Awrite(att, H5T_NATIVE_DOUBLE, buf) >= 0;
}

bool write_attr_str(hid_t loc, const char *name, const char *s, size_t len) {
  std::string buf(len, ' ');
  if (s) {
    const size_t n = std::strlen(s);
    std::memcpy(&buf[0], s, n < len ? n : len);
  }
  H typ(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(typ, len);
  H5Tset_strpad(typ, H5T_STR_SPACEPAD);  // Fortran character semantics
  hsize_t one = 1;
  H spc(H5Screate_simple(1, &one, nullptr), H5Sclose);
  H att(H5Acreate2(loc, name, typ, spc, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return att >= 0 && H5Awrite(att, typ, buf.data()) >= 0;
}

bool read_attr(hid_t loc, const char *name, hid_t type, void *buf) {
  H att(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  return att >= 0 && H5Aread(att, type, buf) >= 0;
}

bool exists(hid_t loc, const char *name) { return H5Lexists(loc, name, H5P_DEFAULT) > 0; }

// datasets (dims in HDF5 order: the Fortran dimensions reversed) ---------
// The division of an array that exceeds the wrapper's chunk limit: the
// HDF5 dimension j (the last Fortran dimension with more than one element)
// and the extent h of the first half, or j = -1 if the array is written
// whole.
int divide(const std::vector<hsize_t> &dims, size_t esize, hsize_t &h) {
  hsize_t n = 1;
  for (hsize_t d : dims) n *= d;
  if (n <= (hsize_t)(g_max_chunk / esize)) return -1;
  for (size_t j = 0; j < dims.size(); j++)
    if (dims[j] > 1) {
      h = (dims[j] - 1) / 2 + 1;  // i - lb + 1 with i = (ub - lb)/2 + lb
      return (int)j;
    }
  return -1;
}

// skip_existing: hdf_rdwr's mode 2 (an existing dataset is left alone),
// applied to each part as the reference applies it after dividing.
bool write_set(hid_t loc, const std::string &name, hid_t type, const void *buf,
               std::vector<hsize_t> dims, bool skip_existing = false) {
  const size_t es = H5Tget_size(type);
  hsize_t h = 0;
  const int j = divide(dims, es, h);
  if (j >= 0) {
    hsize_t inner = 1;
    for (size_t k = j + 1; k < dims.size(); k++) inner *= dims[k];
    std::vector<hsize_t> d0 = dims, d1 = dims;
    d0[j] = h;
    d1[j] = dims[j] - h;
    return write_set(loc, name + "_0", type, buf, d0, skip_existing) &&
           write_set(loc, name + "_1", type, (const char *)buf + es * h * inner, d1,
                     skip_existing);
  }
  if (skip_existing && exists(loc, name.c_str())) return true;
  H spc(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose);
  H set(H5Dcreate2(loc, name.c_str(), type, spc, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Dclose);
  return spc >= 0 && set >= 0 && H5Dwrite(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0;
}

bool read_set(hid_t loc, const std::string &name, hid_t type, void *buf,
              std::vector<hsize_t> dims) {
  const size_t es = H5Tget_size(type);
  hsize_t h = 0;
  const int j = divide(dims, es, h);
  if (j >= 0) {
    hsize_t inner = 1;
    for (size_t k = j + 1; k < dims.size(); k++) inner *= dims[k];
    std::vector<hsize_t> d0 = dims, d1 = dims;
    d0[j] = h;
    d1[j] = dims[j] - h;
    return read_set(loc, name + "_0", type, buf, d0) &&
           read_set(loc, name + "_1", type, (char *)buf + es * h * inner, d1);
  }
  H set(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (set < 0) return false;
  H spc(H5Dget_space(set), H5Sclose);
  hsize_t expect = 1;
  for (hsize_t d : dims) expect *= d;
  if (H5Sget_simple_extent_npoints(spc) != (hssize_t)expect) return false;
  return H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0;
}

int file_version(hid_t file) {
  int version = 0;  // default (correlation.f:1679)
  if (H5Aexists(file, "version") > 0) read_attr(file, "version", H5T_NATIVE_INT, &version);
  return version;
}

std::string map_dir(const char *prefix) { return std::string(prefix) + ".cor.map"; }

bool file_exists(const std::string &f) {
  struct stat st;
  return stat(f.c_str(), &st) == 0;
}

}  // namespace

extern "C" {

int spexcor_last_error(char *buf, size_t len) {
  if (!buf || !len) return 1;
  std::snprintf(buf, len, "%s", g_err.c_str());
  return 0;
}

int spexcor_write(const char *fname, const double kvec[3], int typ, const char *label, int nbasm,
                  const double *coulomb, int nbas, const double *coul, const double *ctrafo,
                  int ncblock, const int *cblock, int oselfc, const double contour[4],
                  int gamma, double plasma, int nfreq, const double *freq, const double *matrix,
                  const double *head, const double *wing, int nfreqc, const double *freqc,
                  const double *matrixc, const double *headc, const double *wingc) {
  if (!fname || !kvec || !contour || typ < 1 || typ > 3 || nbasm <= 0 || nbas <= 0 ||
      nbas > nbasm || nfreq < 0 || nfreqc < 0 || ncblock < 0 || !coulomb || !coul || !ctrafo)
    return fail("spexcor_write: bad arguments");
  if ((nfreq && (!freq || !matrix)) || (nfreqc && (!freqc || !matrixc)) ||
      (ncblock && !cblock))
    return fail("spexcor_write: missing frequency data or cblock");
  if (gamma && ((nfreq && (!head || !wing)) || (nfreqc && (!headc || !wingc))))
    return fail("spexcor_write: the Gamma point needs head and wings");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  // hdf_fopen (Hwrapper.f:81-121): open read-write, create when missing
  FILE *probe = std::fopen(fname, "rb");
  hid_t fid;
  if (probe) {
    std::fclose(probe);
    fid = H5Fopen(fname, H5F_ACC_RDWR, H5P_DEFAULT);
  } else {
    fid = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  }
  H file(fid, H5Fclose);
  if (file < 0) return fail(std::string("spexcor_write: cannot open ") + fname);
  // version 1 (mode 2: written once, otherwise checked)
  const int version = 1;
  if (H5Aexists(file, "version") > 0) {
    int v = -1;
    if (!read_attr(file, "version", H5T_NATIVE_INT, &v) || v != version)
      return fail("spexcor_write: Attribute(s) version incorrect.");
  } else if (!write_attr(file, "version", H5T_NATIVE_INT, &version, 1)) {
    return fail("spexcor_write: cannot write version");
  }
  const std::string gname = group_name(kvec);
  hid_t gid = exists(file, gname.c_str())
                  ? H5Gopen2(file, gname.c_str(), H5P_DEFAULT)
                  : H5Gcreate2(file, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H grp(gid, H5Gclose);
  if (grp < 0) return fail("spexcor_write: cannot open group " + gname);
  // nbasm and the Coulomb matrix (mode 2)
  const hsize_t npm = (hsize_t)nbasm * (nbasm + 1) / 2;
  if (H5Aexists(grp, "nbasm") > 0) {
    int v = -1;
    if (!read_attr(grp, "nbasm", H5T_NATIVE_INT, &v) || v != nbasm)
      return fail("spexcor_write: Attribute(s) nbasm incorrect.");
  } else if (!write_attr(grp, "nbasm", H5T_NATIVE_INT, &nbasm, 1)) {
    return fail("spexcor_write: cannot write nbasm");
  }
  if (!write_set(grp, "coulomb", H5T_NATIVE_DOUBLE, coulomb, mc({npm}), true))
    return fail("spexcor_write: cannot write coulomb");
  // file attributes (mode 2)
  const int nn[2] = {nfreq, nfreqc};
  if (H5Aexists(file, "nfreq,nfreqc") > 0) {
    int v[2];
    if (!read_attr(file, "nfreq,nfreqc", H5T_NATIVE_INT, v) || v[0] != nfreq || v[1] != nfreqc)
      return fail("spexcor_write: Attribute(s) nfreq,nfreqc incorrect.");
  } else {
    if (!write_attr(file, "nfreq,nfreqc", H5T_NATIVE_INT, nn, 2) ||
        (nfreq && !write_attr(file, "freq", H5T_NATIVE_DOUBLE, freq, nfreq)) ||
        (nfreqc && !write_attr_cplx(file, "freqc", freqc, nfreqc)))
      return fail("spexcor_write: cannot write the frequency meshes");
  }
  // an existing matrix set is kept (correlation.f:1860-1866)
  if (H5Aexists(grp, "type") > 0) return fail("spexcor_write: dataset " + gname + " exists");
  const int one_vals[4] = {typ, nbas, ncblock, oselfc};
  const char *one_names[4] = {"type", "dim", "ncblock", "oselfc"};
  if (!write_attr(grp, one_names[0], H5T_NATIVE_INT, &one_vals[0], 1) ||
      !write_attr_str(grp, "label", label, 6))
    return fail("spexcor_write: cannot write the k-point attributes");
  for (int a = 1; a < 4; a++)
    if (!write_attr(grp, one_names[a], H5T_NATIVE_INT, &one_vals[a], 1))
      return fail("spexcor_write: cannot write the k-point attributes");
  if (!write_attr(grp, "contour", H5T_NATIVE_DOUBLE, contour, 4) ||
      (gamma && !write_attr(grp, "plasma", H5T_NATIVE_DOUBLE, &plasma, 1)))
    return fail("spexcor_write: cannot write the k-point attributes");
  const hsize_t nb = nbas, nbm = nbasm, np = (hsize_t)nbas * (nbas + 1) / 2;
  if (!write_set(grp, "coul", H5T_NATIVE_DOUBLE, coul, {nb}) ||
      !write_set(grp, "ctrafo", H5T_NATIVE_DOUBLE, ctrafo, mc({nb, nbm})) ||
      (ncblock && !write_set(grp, "cblock", H5T_NATIVE_INT, cblock, {(hsize_t)ncblock})))
    return fail("spexcor_write: cannot write coul, ctrafo or cblock");
  for (int i = 0; i < nfreq; i++)
    if (!write_set(grp, "matrix_" + std::to_string(i + 1), H5T_NATIVE_DOUBLE,
                   matrix + (g_inv ? 1 : 2) * np * i, mc({np})))
      return fail("spexcor_write: cannot write matrix_" + std::to_string(i + 1));
  for (int i = 0; i < nfreqc; i++)
    if (!write_set(grp, "matrixc_" + std::to_string(i + 1), H5T_NATIVE_DOUBLE,
                   matrixc + 2 * nb * nb * i, {nb, nb, 2}))
      return fail("spexcor_write: cannot write matrixc_" + std::to_string(i + 1));
  if (gamma) {
    const hsize_t nf = nfreq, nfc = nfreqc;
    if ((nfreq && (!write_set(grp, "head", H5T_NATIVE_DOUBLE, head, mc({nf, 6})) ||
                   !write_set(grp, "wing", H5T_NATIVE_DOUBLE, wing, mc({nf, nb, 3})))) ||
        (nfreqc && (!write_set(grp, "headc", H5T_NATIVE_DOUBLE, headc, {nfc, 3, 3, 2}) ||
                    !write_set(grp, "wingc", H5T_NATIVE_DOUBLE, wingc, {nfc, 2, nb, 3, 2}))))
      return fail("spexcor_write: cannot write head or wings");
  }
  return 0;
}

int spexcor_query(const char *fname, const double kvec[3], int *typ, int *nbasm, int *nbas,
                  int *nfreq, int *nfreqc, int *ncblock) {
  if (!fname || !kvec || !typ || !nbasm || !nbas || !nfreq || !nfreqc || !ncblock)
    return fail("spexcor_query: bad arguments");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file < 0) return fail(std::string("File ") + fname + " not found.");
  const int version = file_version(file);
  if (version != 0 && version != 1)
    return fail("COR version number unknown: " + std::to_string(version));
  const std::string gname = group_name(kvec);
  *typ = -1;  // the reference's "nothing for this k point"
  if (!exists(file, gname.c_str())) return 0;
  H grp(H5Gopen2(file, gname.c_str(), H5P_DEFAULT), H5Gclose);
  if (!read_attr(grp, "nbasm", H5T_NATIVE_INT, nbasm))
    return fail("spexcor_query: no nbasm in " + gname);
  *typ = 0;
  if (H5Aexists(grp, "type") <= 0) return 0;
  int nn[2] = {0, 0};
  if (!read_attr(grp, "type", H5T_NATIVE_INT, typ) || !read_attr(grp, "dim", H5T_NATIVE_INT, nbas) ||
      !read_attr(grp, "ncblock", H5T_NATIVE_INT, ncblock) ||
      !read_attr(file, "nfreq,nfreqc", H5T_NATIVE_INT, nn))
    return fail("spexcor_query: incomplete attributes in " + gname);
  *nfreq = nn[0];
  *nfreqc = nn[1];
  return 0;
}

int spexcor_read(const char *fname, const double kvec[3], char *label, double *coulomb,
                 double *coul, double *ctrafo, int *cblock, int *oselfc, double contour[4],
                 double *plasma, double *freq, double *matrix, double *head, double *wing,
                 double *freqc, double *matrixc, double *headc, double *wingc) {
  int typ, nbasm, nbas, nfreq, nfreqc, ncblock;
  if (spexcor_query(fname, kvec, &typ, &nbasm, &nbas, &nfreq, &nfreqc, &ncblock)) return 1;
  if (typ < 1) return fail("spexcor_read: no matrices for this k point");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  const std::string gname = group_name(kvec);
  H grp(H5Gopen2(file, gname.c_str(), H5P_DEFAULT), H5Gclose);
  const hsize_t nb = nbas, np = (hsize_t)nbas * (nbas + 1) / 2,
                npm = (hsize_t)nbasm * (nbasm + 1) / 2;
  const bool gamma = H5Aexists(grp, "plasma") > 0;
  {
    int inv = -1;
    if (spexcor_file_inv(fname, kvec, &inv)) return 1;
    if (inv != (g_inv ? 1 : 0))
      return fail(std::string("spexcor_read: ") + fname + " holds the " +
                  (inv ? "real (INV build)" : "complex") +
                  " layout; select it with spexcor_set_inv");
  }
  if (label) {
    H att(H5Aopen(grp, "label", H5P_DEFAULT), H5Aclose);
    H typ_s(H5Aget_type(att), H5Tclose);
    const size_t len = H5Tget_size(typ_s);
    std::vector<char> buf(len + 1, 0);
    if (H5Aread(att, typ_s, buf.data()) < 0) return fail("spexcor_read: label");
    std::memcpy(label, buf.data(), len < 6 ? len : 6);
  }
  const hsize_t nbm = nbasm;
  if ((coulomb && !read_set(grp, "coulomb", H5T_NATIVE_DOUBLE, coulomb, mc({npm}))) ||
      (coul && !read_set(grp, "coul", H5T_NATIVE_DOUBLE, coul, {nb})) ||
      (ctrafo && !read_set(grp, "ctrafo", H5T_NATIVE_DOUBLE, ctrafo, mc({nb, nbm}))) ||
      (cblock && ncblock && !read_set(grp, "cblock", H5T_NATIVE_INT, cblock, {(hsize_t)ncblock})) ||
      (oselfc && !read_attr(grp, "oselfc", H5T_NATIVE_INT, oselfc)) ||
      (contour && !read_attr(grp, "contour", H5T_NATIVE_DOUBLE, contour)) ||
      (plasma && gamma && !read_attr(grp, "plasma", H5T_NATIVE_DOUBLE, plasma)))
    return fail("spexcor_read: k-point data of " + gname);
  if (nfreq && freq && !read_attr(file, "freq", H5T_NATIVE_DOUBLE, freq))
    return fail("spexcor_read: freq");
  if (nfreqc && freqc && !read_attr(file, "freqc", H5T_NATIVE_DOUBLE, freqc))
    return fail("spexcor_read: freqc");
  const hsize_t nf = nfreq, nfc = nfreqc;
  if (file_version(file) == 0) {
    // one dataset per matrix kind (correlation.f:1799-1807)
    if (matrix && nfreq && !read_set(grp, "matrix", H5T_NATIVE_DOUBLE, matrix, mc({nf, np})))
      return fail("spexcor_read: matrix");
    if (matrixc && nfreqc &&
        !read_set(grp, "matrixc", H5T_NATIVE_DOUBLE, matrixc, {nfc, nb, nb, 2}))
      return fail("spexcor_read: matrixc");
  } else {
    for (int i = 0; matrix && i < nfreq; i++)
      if (!read_set(grp, "matrix_" + std::to_string(i + 1), H5T_NATIVE_DOUBLE,
                    matrix + (g_inv ? 1 : 2) * np * i, mc({np})))
        return fail("spexcor_read: matrix_" + std::to_string(i + 1));
    for (int i = 0; matrixc && i < nfreqc; i++)
      if (!read_set(grp, "matrixc_" + std::to_string(i + 1), H5T_NATIVE_DOUBLE,
                    matrixc + 2 * nb * nb * i, {nb, nb, 2}))
        return fail("spexcor_read: matrixc_" + std::to_string(i + 1));
  }
  if (gamma) {
    if ((nfreq && head && !read_set(grp, "head", H5T_NATIVE_DOUBLE, head, mc({nf, 6}))) ||
        (nfreq && wing && !read_set(grp, "wing", H5T_NATIVE_DOUBLE, wing, mc({nf, nb, 3}))) ||
        (nfreqc && headc && !read_set(grp, "headc", H5T_NATIVE_DOUBLE, headc, {nfc, 3, 3, 2})) ||
        (nfreqc && wingc && !read_set(grp, "wingc", H5T_NATIVE_DOUBLE, wingc, {nfc, 2, nb, 3, 2})))
      return fail("spexcor_read: head or wings");
  }
  return 0;
}

int spexcor_set_inv(int inv) {
  g_inv = inv != 0;
  return 0;
}

int spexcor_file_inv(const char *fname, const double kvec[3], int *inv) {
  if (!fname || !kvec || !inv) return fail("spexcor_file_inv: bad arguments");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file < 0) return fail(std::string("File ") + fname + " not found.");
  const std::string gname = group_name(kvec);
  if (!exists(file, gname.c_str())) return fail("spexcor_file_inv: no group " + gname);
  H grp(H5Gopen2(file, gname.c_str(), H5P_DEFAULT), H5Gclose);
  // the packed Coulomb matrix (or its first part): rank 1 real, rank 2 complex
  std::string name = "coulomb";
  while (!exists(grp, name.c_str())) {
    name += "_0";
    if (name.size() > 64) return fail("spexcor_file_inv: no coulomb dataset in " + gname);
  }
  H set(H5Dopen2(grp, name.c_str(), H5P_DEFAULT), H5Dclose);
  H spc(H5Dget_space(set), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(spc);
  if (rank < 1) return fail("spexcor_file_inv: cannot read " + name);
  hsize_t dims[2] = {0, 0};
  H5Sget_simple_extent_dims(spc, dims, nullptr);
  *inv = (rank == 2 && dims[1] == 2) ? 0 : 1;
  return 0;
}

int spexcor_set_max_chunk(size_t bytes) {
  g_max_chunk = bytes ? bytes : 2147483647;
  return 0;
}

int spexcor_file_name(const char *prefix, int color, char *fname, size_t len) {
  if (!prefix || !fname || !len || color < 0) return fail("spexcor_file_name: bad arguments");
  std::string f = std::string(prefix) + ".cor";
  if (color > 0) f += "." + std::to_string(color);
  if (f.size() + 1 > len) return fail("spexcor_file_name: buffer too small");
  std::memcpy(fname, f.c_str(), f.size() + 1);
  return 0;
}

int spexcor_locate(const char *prefix, const double kvec[3], char *fname, size_t len) {
  if (!prefix || !kvec || !fname || !len) return fail("spexcor_locate: bad arguments");
  int color = 0;
  const std::string dir = map_dir(prefix);
  if (file_exists(dir + "/present")) {
    const std::string entry = dir + "/" + group_name(kvec);
    if (file_exists(entry)) {
      FILE *f = std::fopen(entry.c_str(), "r");
      if (!f || std::fscanf(f, "%d", &color) != 1) {
        if (f) std::fclose(f);
        return fail("Could not read corfile handle.");
      }
      std::fclose(f);
      if (color <= 0) return fail("corfile handle non-positive.");
    }
  }
  return spexcor_file_name(prefix, color, fname, len);
}

int spexcor_map_record(const char *prefix, const double kvec[3], int color) {
  if (!prefix || !kvec || color < 0) return fail("spexcor_map_record: bad arguments");
  const std::string dir = map_dir(prefix), present = dir + "/present";
  const std::string entry = dir + "/" + group_name(kvec);
  if (color > 0) {
    if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST)
      return fail("Could not create directory " + dir);
    if (!file_exists(present)) {
      FILE *f = std::fopen(present.c_str(), "a");
      if (!f) return fail("Could not create empty file in directory " + dir);
      std::fclose(f);
    }
    FILE *f = std::fopen(entry.c_str(), "w");
    if (!f) return fail("Could not write " + entry);
    std::fprintf(f, "%12d\n", color);  // list-directed integer, as write(iunit,*)
    std::fclose(f);
    return 0;
  }
  // a k point written to spex.cor itself: a stale entry is removed
  if (file_exists(present) && file_exists(entry) && unlink(entry.c_str()) != 0)
    return fail("Could not delete " + entry);
  return 0;
}

}  // extern "C"

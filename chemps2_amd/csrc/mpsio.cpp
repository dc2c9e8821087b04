// MPS checkpoints in the reference's HDF5 layout (DMRG::saveMPS, loadDIM, loadMPS,
// DMRGmpsio.cpp:27-162), on top of the public C ABI of libchemps2_amd:
//   /Convergence/Converged_yn              int32[1]
//   /VirtDim_<b>_<N>_<2S>_<I>/Value         int32[1] for every boundary b = 0..L and
//                                           every (N, 2S, I) of SyBookkeeper's ranges
//   /MPS_<site>/Values                      float64[TensorT size], TensorT block order
// Files written here are read by the reference and the other way round.
#include <hdf5.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/chemps2_amd.h"
#include "../../include/chemps2_amd_io.h"

namespace {

struct Range {
  int b, N, S, I;
};

// Every (boundary, N, 2S, irrep) of SyBookkeeper's Nmin/Nmax/TwoSmin/TwoSmax in the
// reference's loop order (SyBookkeeper.cpp:84-92, DMRGmpsio.cpp:47-52).
std::vector<Range> sector_ranges(int L, int N, int TS, int nirrep) {
  std::vector<Range> out;
  for (int b = 0; b <= L; b++) {
    const int nmin = std::max(std::max(0, N + 2 * (b - L)), b - L + (N + TS) / 2);
    const int nmax = std::min(std::min(2 * b, N), b + (N - TS) / 2);
    for (int n = nmin; n <= nmax; n++) {
      const int tmp = L - b - std::abs(N - n - L + b);
      const int smin = std::max(n % 2, TS - tmp), smax = std::min(b - std::abs(b - n), TS + tmp);
      for (int s = smin; s <= smax; s += 2)
        for (int i = 0; i < nirrep; i++) out.push_back({b, n, s, i});
    }
  }
  return out;
}

bool write_int(hid_t file, const std::string& group, const char* name, int value) {
  const hid_t g = H5Gcreate2(file, group.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (g < 0) return false;
  hsize_t one = 1;
  const hid_t sp = H5Screate_simple(1, &one, nullptr);
  const hid_t ds = H5Dcreate2(g, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  const bool ok = ds >= 0 && H5Dwrite(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &value) >= 0;
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  H5Gclose(g);
  return ok;
}

bool read_int(hid_t file, const std::string& group, const char* name, int* value) {
  const hid_t g = H5Gopen2(file, group.c_str(), H5P_DEFAULT);
  if (g < 0) return false;
  const hid_t ds = H5Dopen2(g, name, H5P_DEFAULT);
  const bool ok = ds >= 0 && H5Dread(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) >= 0;
  if (ds >= 0) H5Dclose(ds);
  H5Gclose(g);
  return ok;
}

std::string vdim_group(const Range& r) {
  return "/VirtDim_" + std::to_string(r.b) + "_" + std::to_string(r.N) + "_" + std::to_string(r.S) + "_" + std::to_string(r.I);
}

struct Silence {  // HDF5 prints its error stack by default; errors are returned instead
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Silence() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Silence() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

}  // namespace

extern "C" {

int chm_io_write(const char* path, int L, int N, int two_s, int nirrep, int converged, const int* dims,
                 const double* const* mps, const int64_t* mps_size) {
  if (!path || L < 2 || !dims || !mps || !mps_size) return CHM_ERR_ARG;
  Silence quiet;
  const std::vector<Range> rs = sector_ranges(L, N, two_s, nirrep);
  const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return CHM_ERR_ARG;
  bool ok = write_int(f, "/Convergence", "Converged_yn", converged ? 1 : 0);
  for (size_t q = 0; ok && q < rs.size(); q++) ok = write_int(f, vdim_group(rs[q]), "Value", dims[q]);
  for (int s = 0; ok && s < L; s++) {
    const hid_t g = H5Gcreate2(f, ("/MPS_" + std::to_string(s)).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) { ok = false; break; }
    hsize_t n = (hsize_t)mps_size[s];
    const hid_t sp = H5Screate_simple(1, &n, nullptr);
    const hid_t ds = H5Dcreate2(g, "Values", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    ok = ds >= 0 && (n == 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, mps[s]) >= 0);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    H5Gclose(g);
  }
  H5Fclose(f);
  return ok ? CHM_OK : CHM_ERR_STATE;
}

int chm_io_read_dims(const char* path, int L, int N, int two_s, int nirrep, int* dims, int* n_dims) {
  if (!path || L < 2 || !n_dims) return CHM_ERR_ARG;
  const std::vector<Range> rs = sector_ranges(L, N, two_s, nirrep);
  if (!dims) { *n_dims = (int)rs.size(); return CHM_OK; }
  if (*n_dims < (int)rs.size()) return CHM_ERR_SIZE;
  Silence quiet;
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return CHM_ERR_ARG;
  bool ok = true;
  for (size_t q = 0; ok && q < rs.size(); q++) ok = read_int(f, vdim_group(rs[q]), "Value", &dims[q]);
  H5Fclose(f);
  *n_dims = (int)rs.size();
  return ok ? CHM_OK : CHM_ERR_MISSING;
}

int chm_io_read_site(const char* path, int site, double* values, int64_t size, int* converged) {
  if (!path || site < 0 || (size > 0 && !values)) return CHM_ERR_ARG;
  Silence quiet;
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return CHM_ERR_ARG;
  bool ok = true;
  if (converged) ok = read_int(f, "/Convergence", "Converged_yn", converged);
  const hid_t g = ok ? H5Gopen2(f, ("/MPS_" + std::to_string(site)).c_str(), H5P_DEFAULT) : -1;
  int rc = CHM_OK;
  if (g < 0) rc = CHM_ERR_MISSING;
  else {
    const hid_t ds = H5Dopen2(g, "Values", H5P_DEFAULT);
    if (ds < 0) rc = CHM_ERR_MISSING;
    else {
      const hid_t sp = H5Dget_space(ds);
      hsize_t n = 0;
      H5Sget_simple_extent_dims(sp, &n, nullptr);
      if ((int64_t)n != size) rc = CHM_ERR_SIZE;
      else if (n > 0 && H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, values) < 0) rc = CHM_ERR_STATE;
      H5Sclose(sp);
      H5Dclose(ds);
    }
    H5Gclose(g);
  }
  H5Fclose(f);
  return rc;
}

int chm_mps_save(chm_ctx* ctx, const char* path, int converged) {
  int L, N, TS, irrep, nirrep;
  int rc = chm_get_problem(ctx, &L, &N, &TS, &irrep, &nirrep);
  if (rc != CHM_OK) return rc;
  const std::vector<Range> rs = sector_ranges(L, N, TS, nirrep);
  std::vector<int> dims(rs.size(), 0);
  for (int b = 0; b <= L; b++) {
    int n = 0;
    if ((rc = chm_get_sectors(ctx, b, &n, nullptr, nullptr, nullptr, nullptr)) != CHM_OK) return rc;
    std::vector<int> Ns(n), Ss(n), Is(n), ds(n);
    if ((rc = chm_get_sectors(ctx, b, &n, Ns.data(), Ss.data(), Is.data(), ds.data())) != CHM_OK) return rc;
    for (int q = 0; q < n; q++)
      for (size_t r = 0; r < rs.size(); r++)
        if (rs[r].b == b && rs[r].N == Ns[q] && rs[r].S == Ss[q] && rs[r].I == Is[q]) dims[r] = ds[q];
  }
  std::vector<std::vector<double>> data(L);
  std::vector<const double*> ptr(L);
  std::vector<int64_t> size(L);
  for (int s = 0; s < L; s++) {
    int64_t n = 0;
    if ((rc = chm_mps_layout_size(ctx, s, &n)) != CHM_OK) return rc;
    data[s].resize(std::max<int64_t>(n, 1));
    if ((rc = chm_mps_download(ctx, s, data[s].data(), n)) != CHM_OK) return rc;
    ptr[s] = data[s].data();
    size[s] = n;
  }
  return chm_io_write(path, L, N, TS, nirrep, converged, dims.data(), ptr.data(), size.data());
}

int chm_mps_load_dims(chm_ctx* ctx, const char* path) {
  int L, N, TS, irrep, nirrep;
  int rc = chm_get_problem(ctx, &L, &N, &TS, &irrep, &nirrep);
  if (rc != CHM_OK) return rc;
  const std::vector<Range> rs = sector_ranges(L, N, TS, nirrep);
  std::vector<int> dims(rs.size(), 0);
  int n = (int)rs.size();
  if ((rc = chm_io_read_dims(path, L, N, TS, nirrep, dims.data(), &n)) != CHM_OK) return rc;
  for (int b = 0; b <= L; b++) {
    std::vector<int> Ns, Ss, Is, ds;
    for (size_t r = 0; r < rs.size(); r++)
      if (rs[r].b == b && dims[r] > 0) { Ns.push_back(rs[r].N); Ss.push_back(rs[r].S); Is.push_back(rs[r].I); ds.push_back(dims[r]); }
    if ((rc = chm_set_sectors(ctx, b, (int)Ns.size(), Ns.data(), Ss.data(), Is.data(), ds.data())) != CHM_OK) return rc;
  }
  return CHM_OK;
}

int chm_mps_load(chm_ctx* ctx, const char* path, int* converged) {
  int L, N, TS, irrep, nirrep;
  int rc = chm_get_problem(ctx, &L, &N, &TS, &irrep, &nirrep);
  if (rc != CHM_OK) return rc;
  for (int s = 0; s < L; s++) {
    int64_t n = 0;
    if ((rc = chm_mps_layout_size(ctx, s, &n)) != CHM_OK) return rc;
    std::vector<double> v(std::max<int64_t>(n, 1));
    if ((rc = chm_io_read_site(path, s, v.data(), n, s == 0 ? converged : nullptr)) != CHM_OK) return rc;
    if ((rc = chm_mps_upload(ctx, s, v.data(), n)) != CHM_OK) return rc;
  }
  return CHM_OK;
}

}  // extern "C"

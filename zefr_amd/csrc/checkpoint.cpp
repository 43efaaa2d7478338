// checkpoint.cpp -- PyFR-format solution files (.pyfrs): write and restart.
//
// FRSolver::write_solution_pyfr (src/solver.cpp:4275-4668) and FRSolver::restart_pyfr
// (src/solver.cpp:4670-4837) for one element block on one rank.  A .pyfrs file is HDF5
// with three scalar string datasets ("config", "stats", "mesh_uuid") and one double
// dataset per element type, "soln_<quad|hex>_p0", shaped [nSpts][nVars][nEles].  The
// HDF5 C library is loaded at run time (dlopen), so the solver library has no link-time
// dependency on it; a missing library is reported when a checkpoint is requested.
// Deviations from the reference: the stats record tcurr with 17 significant digits (the
// reference streams it at the default 6, so a restart loses time precision), and the
// config records the deck's gamma (the reference always writes 1.4).
#include <dlfcn.h>
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

namespace {

// The slice of the HDF5 1.10 C API used here (H5public.h, H5Ipublic.h: hid_t is 64-bit).
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;  // H5F_ACC_TRUNC / _RDONLY
constexpr hid_t kDefault = 0;                                    // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr int kScalar = 0;                                       // H5S_SCALAR
constexpr size_t kVariable = (size_t)-1;                         // H5T_VARIABLE

struct H5 {
  void *so = nullptr;
  herr_t (*open)(void);
  herr_t (*eset_auto)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*tcopy)(hid_t);
  herr_t (*tset_size)(hid_t, size_t);
  size_t (*tget_size)(hid_t);
  htri_t (*tis_variable_str)(hid_t);
  herr_t (*tclose)(hid_t);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_space)(hid_t);
  hid_t (*dget_type)(hid_t);
  herr_t (*free_memory)(void *);
  hid_t *c_s1 = nullptr, *native_double = nullptr;  // H5T_C_S1_g, H5T_NATIVE_DOUBLE_g

  template <typename F>
  void sym(F &f, const char *name) {
    f = reinterpret_cast<F>(dlsym(so, name));
    if (!f) raise(std::string("HDF5 library lacks ") + name);
  }
};

H5 &h5() {
  static H5 lib;
  if (lib.so) return lib;
  std::vector<std::string> names;
  if (const char *env = std::getenv("ZEFR_HDF5_LIBRARY")) names.push_back(env);
  names.push_back("libhdf5.so");
  names.push_back("/opt/conda/lib/libhdf5.so");
  std::string tried;
  for (auto &n : names) {
    lib.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (lib.so) break;
    tried += " " + n;
  }
  if (!lib.so)
    raise("PyFR solution files need the HDF5 C library; none found (tried" + tried +
          "); set ZEFR_HDF5_LIBRARY to its path");
  try {
    lib.sym(lib.open, "H5open");
    lib.sym(lib.eset_auto, "H5Eset_auto2");
    lib.sym(lib.fcreate, "H5Fcreate");
    lib.sym(lib.fopen, "H5Fopen");
    lib.sym(lib.fclose, "H5Fclose");
    lib.sym(lib.screate, "H5Screate");
    lib.sym(lib.screate_simple, "H5Screate_simple");
    lib.sym(lib.sclose, "H5Sclose");
    lib.sym(lib.sget_ndims, "H5Sget_simple_extent_ndims");
    lib.sym(lib.sget_dims, "H5Sget_simple_extent_dims");
    lib.sym(lib.tcopy, "H5Tcopy");
    lib.sym(lib.tset_size, "H5Tset_size");
    lib.sym(lib.tget_size, "H5Tget_size");
    lib.sym(lib.tis_variable_str, "H5Tis_variable_str");
    lib.sym(lib.tclose, "H5Tclose");
    lib.sym(lib.dcreate, "H5Dcreate2");
    lib.sym(lib.dopen, "H5Dopen2");
    lib.sym(lib.dwrite, "H5Dwrite");
    lib.sym(lib.dread, "H5Dread");
    lib.sym(lib.dclose, "H5Dclose");
    lib.sym(lib.dget_space, "H5Dget_space");
    lib.sym(lib.dget_type, "H5Dget_type");
    lib.sym(lib.free_memory, "H5free_memory");
    lib.c_s1 = static_cast<hid_t *>(dlsym(lib.so, "H5T_C_S1_g"));
    lib.native_double = static_cast<hid_t *>(dlsym(lib.so, "H5T_NATIVE_DOUBLE_g"));
    if (!lib.c_s1 || !lib.native_double) raise("HDF5 library lacks its predefined type ids");
    if (lib.open() < 0) raise("H5open failed");
    lib.eset_auto(kDefault, nullptr, nullptr);  // errors come back as return codes
  } catch (...) {
    dlclose(lib.so);
    lib.so = nullptr;
    throw;
  }
  return lib;
}

void h5check(long long rc, const std::string &what) {
  if (rc < 0) raise("HDF5: " + what + " failed");
}

void write_string(H5 &h, hid_t file, const char *name, const std::string &value) {
  const hid_t st = h.tcopy(*h.c_s1);
  h5check(st, "H5Tcopy");
  h5check(h.tset_size(st, kVariable), "H5Tset_size");
  const hid_t sp = h.screate(kScalar);
  const hid_t ds = h.dcreate(file, name, st, sp, kDefault, kDefault, kDefault);
  const char *p = value.c_str();
  const herr_t rc = ds < 0 ? -1 : h.dwrite(ds, st, kDefault, kDefault, kDefault, &p);
  if (ds >= 0) h.dclose(ds);
  h.sclose(sp);
  h.tclose(st);
  h5check(rc, std::string("writing dataset ") + name);
}

std::string read_string(H5 &h, hid_t file, const char *name) {
  const hid_t ds = h.dopen(file, name, kDefault);
  if (ds < 0) raise(std::string("PyFR file has no dataset '") + name + "'");
  const hid_t ft = h.dget_type(ds);
  std::string out;
  herr_t rc = 0;
  if (h.tis_variable_str(ft) > 0) {
    const hid_t mt = h.tcopy(*h.c_s1);
    h.tset_size(mt, kVariable);
    char *p = nullptr;
    rc = h.dread(ds, mt, kDefault, kDefault, kDefault, &p);
    if (rc >= 0 && p) {
      out = p;
      h.free_memory(p);
    }
    h.tclose(mt);
  } else {  // fixed-length string
    const size_t n = h.tget_size(ft);
    std::vector<char> buf(n + 1, '\0');
    rc = h.dread(ds, ft, kDefault, kDefault, kDefault, buf.data());
    out.assign(buf.data());
  }
  h.tclose(ft);
  h.dclose(ds);
  h5check(rc, std::string("reading dataset ") + name);
  return out;
}

// hash_str (src/funcs.cpp:1536-1545, djb2 over the bytes as signed chars) of the whole
// mesh file, as generate_mesh_id (src/funcs.cpp:1547-1555) makes the mesh_uuid.
std::string mesh_uuid(const std::string &meshfile) {
  std::ifstream f(meshfile, std::ios::binary);
  if (!f.is_open()) raise("cannot open mesh file " + meshfile);
  std::ostringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  unsigned long h = 5381;
  for (char ch : s) {
    if (ch == '\0') break;
    h = ((h << 5) + h) + (long)(signed char)ch;
  }
  return std::to_string(h);
}

const char *etype_name(int etype) { return etype == QUAD_T ? "quad" : "hex"; }

std::string fmt17(double x) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", x);
  return b;
}

}  // namespace

// Host copy of the current solution [E][V][nS]: from the device when it is set up.
static void sync_host_state(FRSolver &s) {
  Elements &e = *s.eles;
  e.U_host.resize(e.state_size());
  if (!s.device_ready) return;
  hip_or_throw(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  hip_or_throw(hipMemcpy(e.U_host.data(), e.U_spts_d.p, e.state_size() * sizeof(double), hipMemcpyDeviceToHost),
               "D2H U_spts");
}

std::string FRSolver::write_solution_pyfr(const std::string &prefix) {
  H5 &h = h5();
  Elements &e = *eles;
  sync_host_state(*this);
  unsigned iter = (unsigned)current_iter;
  if (input->p_multi && input->nLevels > 0 && input->mg_steps[0] > 0) iter /= input->mg_steps[0];
  const std::string dir = input->output_prefix;
  if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) raise("cannot create output directory " + dir);
  const std::string filename = dir + "/" + prefix + "-" + std::to_string(iter) + ".pyfrs";

  std::ostringstream cfg;
  cfg << "[constants]\ngamma = " << fmt17(input->gamma) << "\n\n";
  cfg << "[solver]\nsystem = " << (input->viscous ? "navier-stokes" : "euler") << "\n";
  cfg << "order = " << order << "\n\n";
  if (e.nDims == 2) {
    cfg << "[solver-elements-tri]\nsoln-pts = williams-shunn\n\n";
    cfg << "[solver-elements-quad]\nsoln-pts = gauss-legendre\n\n";
  } else {
    cfg << "[solver-elements-tet]\nsoln-pts = shunn-ham\n\n";
    cfg << "[solver-elements-pri]\nsoln-pts = williams-shunn~gauss-legendre\n\n";
    cfg << "[solver-elements-hex]\nsoln-pts = gauss-legendre\n\n";
  }
  std::ostringstream st;
  st << "[data]\nfields = rho,rhou,rhov," << (e.nDims == 3 ? "rhoW," : "") << "E\nprefix = soln\n\n";
  st << "[solver-time-integrator]\ntcurr = " << fmt17(flow_time) << "\nnsteps = " << current_iter
     << "\nnrjctsteps = 0\n\n";

  // [E][V][nS] -> [nS][V][E]
  const int E = e.nEles, V = e.nVars, nS = e.nSpts;
  std::vector<double> data((size_t)E * V * nS);
  for (int el = 0; el < E; el++)
    for (int v = 0; v < V; v++)
      for (int s = 0; s < nS; s++) data[((size_t)s * V + v) * E + el] = e.U_host[((size_t)el * V + v) * nS + s];

  const hid_t file = h.fcreate(filename.c_str(), kAccTrunc, kDefault, kDefault);
  if (file < 0) raise("cannot create " + filename);
  try {
    write_string(h, file, "config", cfg.str());
    write_string(h, file, "stats", st.str());
    write_string(h, file, "mesh_uuid", mesh_uuid(input->meshfile));
    const hsize_t dims[3] = {(hsize_t)nS, (hsize_t)V, (hsize_t)E};
    const hid_t sp = h.screate_simple(3, dims, nullptr);
    const std::string name = std::string("soln_") + etype_name(e.etype) + "_p0";
    const hid_t ds = h.dcreate(file, name.c_str(), *h.native_double, sp, kDefault, kDefault, kDefault);
    const herr_t rc = ds < 0 ? -1 : h.dwrite(ds, *h.native_double, kDefault, kDefault, kDefault, data.data());
    if (ds >= 0) h.dclose(ds);
    h.sclose(sp);
    h5check(rc, "writing dataset " + name);
  } catch (...) {
    h.fclose(file);
    throw;
  }
  h5check(h.fclose(file), "closing " + filename);
  return filename;
}

void FRSolver::restart_pyfr(const std::string &restart_case, unsigned restart_iter) {
  H5 &h = h5();
  Elements &e = *eles;
  const std::string filename = restart_case + "/" + restart_case + "-" + std::to_string(restart_iter) + ".pyfrs";
  current_iter = (int)restart_iter;
  input->iter = restart_iter;
  input->initIter = restart_iter;
  const hid_t file = h.fopen(filename.c_str(), kAccRdonly, kDefault);
  if (file < 0) raise("cannot open restart file " + filename);
  std::vector<double> data;
  hsize_t dims[3] = {0, 0, 0};
  std::string stats;
  try {
    if (read_string(h, file, "mesh_uuid") != mesh_uuid(input->meshfile))
      raise("Restart Error - Mesh and solution files do not mesh [mesh_uuid].");
    (void)read_string(h, file, "config");
    stats = read_string(h, file, "stats");
    const std::string name = std::string("soln_") + etype_name(e.etype) + "_p0";
    const hid_t ds = h.dopen(file, name.c_str(), kDefault);
    if (ds < 0) raise("restart file has no dataset " + name);
    const hid_t sp = h.dget_space(ds);
    const int rank = h.sget_ndims(sp);
    if (rank == 3) h.sget_dims(sp, dims, nullptr);
    h.sclose(sp);
    if (rank != 3) {
      h.dclose(ds);
      raise("Improper DataSpace rank for solution data.");
    }
    if (dims[2] != (hsize_t)e.nEles || dims[1] != (hsize_t)e.nVars) {
      h.dclose(ds);
      raise("Size of solution data set does not match that from mesh.");
    }
    data.resize((size_t)(dims[0] * dims[1] * dims[2]));
    const herr_t rc = h.dread(ds, *h.native_double, kDefault, kDefault, kDefault, data.data());
    h.dclose(ds);
    h5check(rc, "reading dataset " + name);
  } catch (...) {
    h.fclose(file);
    throw;
  }
  h.fclose(file);

  // process_restart_stats (src/solver.cpp:4839-4897): tcurr
  std::istringstream ss(stats);
  std::string line;
  while (std::getline(ss, line)) {
    std::istringstream ls(line);
    std::string key, eq;
    double t;
    if ((ls >> key) && key == "tcurr" && (ls >> eq >> t)) {
      flow_time = t;
      input->time = t;
    }
  }

  const int E = e.nEles, V = e.nVars, nS = e.nSpts, D = e.nDims;
  const int nR = (int)dims[0];
  e.U_host.assign((size_t)E * V * nS, 0.0);
  if (nR == nS) {
    for (int el = 0; el < E; el++)
      for (int v = 0; v < V; v++)
        for (int s = 0; s < nS; s++) e.U_host[((size_t)el * V + v) * nS + s] = data[((size_t)s * V + v) * E + el];
  } else {
    // a solution of another order: interpolate from its Gauss-Legendre points
    // (set_oppRestart with use_shape false, src/quads.cpp:209-237, src/hexas.cpp:304-335)
    const int n1 = (int)std::lround(std::pow((double)nR, 1.0 / D));
    int cube = 1;
    for (int d = 0; d < D; d++) cube *= n1;
    if (cube != nR || n1 < 1 || n1 > 10) raise("restart file: point count does not match a quad/hex order 0..9");
    const auto xr = gauss_legendre_points((unsigned)n1);
    const auto xs = gauss_legendre_points((unsigned)(order + 1));
    const int N = order + 1;
    std::vector<double> L1((size_t)N * n1);  // L1[a][i] = l_i(xs[a])
    for (int a = 0; a < N; a++)
      for (int i = 0; i < n1; i++) L1[(size_t)a * n1 + i] = lagrange(xr, (unsigned)i, xs[a]);
    std::vector<double> opp((size_t)nS * nR);
    for (int s = 0; s < nS; s++) {
      const int a0 = s % N, a1 = (s / N) % N, a2 = s / (N * N);
      for (int r = 0; r < nR; r++) {
        const int i0 = r % n1, i1 = (r / n1) % n1, i2 = r / (n1 * n1);
        double w = L1[(size_t)a0 * n1 + i0] * L1[(size_t)a1 * n1 + i1];
        if (D == 3) w *= L1[(size_t)a2 * n1 + i2];
        opp[(size_t)s * nR + r] = w;
      }
    }
    for (int el = 0; el < E; el++)
      for (int v = 0; v < V; v++)
        for (int s = 0; s < nS; s++) {
          double acc = 0.0;
          for (int r = 0; r < nR; r++) acc += opp[(size_t)s * nR + r] * data[((size_t)r * V + v) * E + el];
          e.U_host[((size_t)el * V + v) * nS + s] = acc;
        }
  }
  if (device_ready) {
    hip_or_throw(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_or_throw(hipMemcpy(e.U_spts_d.p, e.U_host.data(), e.state_size() * sizeof(double), hipMemcpyHostToDevice),
                 "H2D U_spts");
  }
}

}  // namespace zfr

// pybind11 front end: the reference's Python class `Molecule`
// (src/pybind.cpp:9-305, host façade src/pymolecule.cu:18-233) over the C ABI.
//
// Both spellings of the density methods are exported: the ones the reference
// binds (compute_density, compute_gradient, compute_hessian,
// compute_laplacian, src/pybind.cpp:22-74) and the ones its Python tests call
// (compute_electron_density, ..., tests/test_density.py:21).  Points are an
// (N, 3) array of Cartesian coordinates in bohr, row-major (the reference's
// descriptor methods pass row-major points to column-major readers,
// src/pymolecule.cu:141-233; here every method reads them as (N, 3)).
// The GIL is released while the GPU works.  An optional `precision`
// argument selects "fp64" (default, parity with the reference) or "mixed".

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "chemtools_amd.h"

namespace py = pybind11;

namespace {

using Points = py::array_t<double, py::array::c_style | py::array::forcecast>;

void check(int status) {
  if (status != CT_OK) throw std::runtime_error(std::string("chemtools_amd: ") + ct_last_error());
}

int parse_precision(const std::string& p) {
  if (p == "fp64" || p == "double" || p == "float64") return CT_PREC_FP64;
  if (p == "mixed") return CT_PREC_MIXED;
  throw std::invalid_argument("precision must be 'fp64' or 'mixed'");
}

int64_t n_points(const Points& pts) {
  if (pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("points must have shape (N, 3)");
  return (int64_t)pts.shape(0);
}

class Molecule {
 public:
  explicit Molecule(const std::string& path) : path_(path) { check(ct_mol_create_from_file(path.c_str(), &h_)); }

  Molecule(py::array_t<int32_t, py::array::c_style | py::array::forcecast> shell_atom,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> shell_l,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> shell_nprim,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> shell_prim_off,
           py::array_t<double, py::array::c_style | py::array::forcecast> exps,
           py::array_t<double, py::array::c_style | py::array::forcecast> coefs,
           py::array_t<double, py::array::c_style | py::array::forcecast> atom_xyz,
           py::array_t<double, py::array::c_style | py::array::forcecast> charges,
           py::array_t<double, py::array::c_style | py::array::forcecast> P, py::object numbers) {
    const int nshell = (int)shell_l.size();
    if (shell_atom.size() != nshell || shell_nprim.size() != nshell || shell_prim_off.size() != nshell)
      throw std::invalid_argument("shell arrays must have equal length");
    if (exps.size() != coefs.size()) throw std::invalid_argument("exps and coefs must have equal length");
    if (atom_xyz.ndim() != 2 || atom_xyz.shape(1) != 3) throw std::invalid_argument("atom_xyz must be (natom, 3)");
    const int natom = (int)atom_xyz.shape(0);
    if (charges.size() != natom) throw std::invalid_argument("charges must have natom entries");
    if (P.ndim() != 2 || P.shape(0) != P.shape(1)) throw std::invalid_argument("P must be square");
    ct_basis_desc d{nshell, shell_atom.data(), shell_l.data(), shell_nprim.data(), shell_prim_off.data(),
                    exps.data(), coefs.data()};
    check(ct_mol_create(&d, atom_xyz.data(), charges.data(), natom, P.data(), (int32_t)P.shape(0), &h_));
    if (!numbers.is_none()) {
      auto nums = numbers.cast<py::array_t<int32_t, py::array::c_style | py::array::forcecast>>();
      numbers_override_.assign(nums.data(), nums.data() + nums.size());
    }
  }

  ~Molecule() {
    if (h_) ct_mol_destroy(h_);
  }

  int nbasis() const { return ct_mol_nbasis(h_); }
  int natom() const { return ct_mol_natom(h_); }
  std::string file_path() const { return path_; }

  py::array_t<double> coordinates() const {
    py::array_t<double> out({natom(), 3});
    check(ct_mol_get_coordinates(h_, out.mutable_data()));
    return out;
  }
  py::array_t<int32_t> numbers() const {
    py::array_t<int32_t> out(natom());
    if (!numbers_override_.empty()) {
      std::memcpy(out.mutable_data(), numbers_override_.data(), sizeof(int32_t) * numbers_override_.size());
      return out;
    }
    check(ct_mol_get_numbers(h_, out.mutable_data()));
    return out;
  }
  py::array_t<double> charges() const {
    py::array_t<double> out(natom());
    check(ct_mol_get_charges(h_, out.mutable_data()));
    return out;
  }
  py::array_t<double> density_matrix() const {
    py::array_t<double> out({nbasis(), nbasis()});
    check(ct_mol_get_density_matrix(h_, out.mutable_data()));
    return out;
  }
  void set_coordinates(py::array_t<double, py::array::c_style | py::array::forcecast> xyz) {
    if (xyz.size() != 3 * natom()) throw std::invalid_argument("coordinates must be (natom, 3)");
    check(ct_mol_set_coordinates(h_, xyz.data()));
  }
  void set_density_matrix(py::array_t<double, py::array::c_style | py::array::forcecast> P) {
    if (P.size() != (py::ssize_t)nbasis() * nbasis()) throw std::invalid_argument("P must be (nbasis, nbasis)");
    check(ct_mol_set_density_matrix(h_, P.data()));
  }
  void set_workspace_limit(int64_t bytes) { check(ct_mol_set_workspace_limit(h_, bytes)); }
  void set_screening(double eps) { check(ct_mol_set_screening(h_, eps)); }

  // Generic multi-quantity evaluation: returns {name: array}.
  py::dict compute(const Points& pts, const std::vector<std::string>& names, const std::string& precision,
                   double a, double b, double alpha, double mixed_tol, const std::string& route) {
    const int64_t n = n_points(pts);
    return compute_impl(pts.data(), nullptr, {n}, names, precision, a, b, alpha, mixed_tol, route);
  }

  // The same on a uniform grid generated on the device: point (i, j, k) =
  // origin + i axes[0] + j axes[1] + k axes[2]; arrays of shape (n0, n1, n2, ...).
  py::dict compute_on_cubic_grid(py::array_t<double, py::array::c_style | py::array::forcecast> origin,
                                 py::array_t<double, py::array::c_style | py::array::forcecast> axes,
                                 py::array_t<int64_t, py::array::c_style | py::array::forcecast> shape,
                                 const std::vector<std::string>& names, const std::string& precision, double a,
                                 double b, double alpha, double mixed_tol, const std::string& route) {
    if (origin.size() != 3 || axes.size() != 9 || shape.size() != 3)
      throw std::invalid_argument("origin must have 3 entries, axes 3 x 3, shape 3");
    Cube c;
    std::memcpy(c.origin, origin.data(), sizeof(c.origin));
    std::memcpy(c.axes, axes.data(), sizeof(c.axes));
    std::memcpy(c.shape, shape.data(), sizeof(c.shape));
    for (int k = 0; k < 3; ++k)
      if (c.shape[k] <= 0) throw std::invalid_argument("grid shape must be positive");
    return compute_impl(nullptr, &c, {c.shape[0], c.shape[1], c.shape[2]}, names, precision, a, b, alpha,
                        mixed_tol, route);
  }

  // compute_electron_density_cubic (src/pymolecule.cu:59-71); `disp` (progress
  // printing in the reference) is accepted and ignored.
  py::array_t<double> density_cubic(py::array_t<double, py::array::c_style | py::array::forcecast> origin,
                                    py::array_t<double, py::array::c_style | py::array::forcecast> axes,
                                    py::array_t<int64_t, py::array::c_style | py::array::forcecast> shape, bool,
                                    const std::string& precision) {
    py::dict d = compute_on_cubic_grid(origin, axes, shape, {"rho"}, precision, 0, 0, 0, 0, "auto");
    return d["rho"].cast<py::array_t<double>>();
  }
  py::array_t<double> one(const Points& pts, uint32_t bit, std::vector<py::ssize_t> extra,
                          const std::string& precision, double a = 0, double b = 0, double alpha = 0) {
    const int64_t n = n_points(pts);
    std::vector<py::ssize_t> shape{n};
    for (auto e : extra) shape.push_back(e);
    py::array_t<double> arr(shape);
    ct_outputs out;
    std::memset(&out, 0, sizeof(out));
    double* p = arr.mutable_data();
    switch (bit) {
      case CT_Q_RHO: out.rho = p; break;
      case CT_Q_GRAD: out.grad = p; break;
      case CT_Q_HESS: out.hess = p; break;
      case CT_Q_LAP: out.lap = p; break;
      case CT_Q_TAU_PD: out.tau_pd = p; break;
      case CT_Q_RDG: out.rdg = p; break;
      case CT_Q_SHANNON: out.shannon = p; break;
      case CT_Q_WEIZSACKER: out.weizsacker = p; break;
      case CT_Q_THOMAS_FERMI: out.thomas_fermi = p; break;
      case CT_Q_GGE: out.gge = p; break;
      case CT_Q_GENERAL_KED: out.general_ked = p; break;
      case CT_Q_KED_ALPHA: out.ked_alpha = p; break;
    }
    ct_eval_params par{};
    par.gge_a = a;
    par.gge_b = b;
    par.ked_a = alpha;
    par.ked_alpha = alpha;
    ct_eval_report rep{};
    const int prec = parse_precision(precision);
    int st;
    {
      py::gil_scoped_release nogil;
      st = ct_eval(h_, pts.data(), n, CT_LAYOUT_ROW, bit, prec, &par, &out, &rep);
    }
    check(st);
    last_report_ = rep;
    return arr;
  }

  py::array_t<double> esp(const Points& pts, const std::string& precision) {
    const int64_t n = n_points(pts);
    py::array_t<double> arr(n);
    const int prec = parse_precision(precision);
    int st;
    {
      py::gil_scoped_release nogil;
      st = ct_esp(h_, pts.data(), n, CT_LAYOUT_ROW, prec, arr.mutable_data());
    }
    check(st);
    return arr;
  }

  py::tuple backward(const Points& pts, py::array_t<double, py::array::c_style | py::array::forcecast> w,
                     const std::string& precision) {
    const int64_t n = n_points(pts);
    if (w.size() != n) throw std::invalid_argument("dL/drho must have N entries");
    py::array_t<double> dP({nbasis(), nbasis()});
    py::array_t<double> dR({natom(), 3});
    const int prec = parse_precision(precision);
    int st;
    {
      py::gil_scoped_release nogil;
      st = ct_backward(h_, pts.data(), n, CT_LAYOUT_ROW, w.data(), prec, dP.mutable_data(), dR.mutable_data());
    }
    check(st);
    return py::make_tuple(dP, dR);
  }

  py::dict last_report() const {
    py::dict d;
    d["n_points"] = last_report_.n_points;
    d["n_reevaluated"] = last_report_.n_reevaluated;
    d["n_chunks"] = last_report_.n_chunks;
    d["max_error_estimate"] = last_report_.max_error_estimate;
    d["factor_rank"] = last_report_.factor_rank;
    d["active_fraction"] = last_report_.active_fraction;
    return d;
  }

  uintptr_t handle() const { return reinterpret_cast<uintptr_t>(h_); }

 private:
  struct Cube {
    double origin[3];
    double axes[9];
    int64_t shape[3];
  };

  py::dict compute_impl(const double* pts, const Cube* cube, std::vector<py::ssize_t> lead,
                        const std::vector<std::string>& names, const std::string& precision, double a, double b,
                        double alpha, double mixed_tol, const std::string& route) {
    static const std::map<std::string, uint32_t> bits = {
        {"rho", CT_Q_RHO},         {"grad", CT_Q_GRAD},
        {"hess", CT_Q_HESS},       {"lap", CT_Q_LAP},
        {"tau_pd", CT_Q_TAU_PD},   {"rdg", CT_Q_RDG},
        {"shannon", CT_Q_SHANNON}, {"weizsacker", CT_Q_WEIZSACKER},
        {"thomas_fermi", CT_Q_THOMAS_FERMI}, {"gge", CT_Q_GGE},
        {"general_ked", CT_Q_GENERAL_KED},   {"ked_alpha", CT_Q_KED_ALPHA}};
    int64_t n = 1;
    for (auto e : lead) n *= e;
    uint32_t q = 0;
    for (const auto& nm : names) {
      auto it = bits.find(nm);
      if (it == bits.end()) throw std::invalid_argument("unknown quantity '" + nm + "'");
      q |= it->second;
    }
    ct_outputs out;
    std::memset(&out, 0, sizeof(out));
    py::dict res;
    auto make = [&](const char* nm, uint32_t bit, double*& slot, std::vector<py::ssize_t> tail) {
      if (!(q & bit)) return;
      std::vector<py::ssize_t> shape = lead;
      shape.insert(shape.end(), tail.begin(), tail.end());
      py::array_t<double> arr(shape);
      slot = arr.mutable_data();
      res[nm] = arr;
    };
    make("rho", CT_Q_RHO, out.rho, {});
    make("grad", CT_Q_GRAD, out.grad, {3});
    make("hess", CT_Q_HESS, out.hess, {3, 3});
    make("lap", CT_Q_LAP, out.lap, {});
    make("tau_pd", CT_Q_TAU_PD, out.tau_pd, {});
    make("rdg", CT_Q_RDG, out.rdg, {});
    make("shannon", CT_Q_SHANNON, out.shannon, {});
    make("weizsacker", CT_Q_WEIZSACKER, out.weizsacker, {});
    make("thomas_fermi", CT_Q_THOMAS_FERMI, out.thomas_fermi, {});
    make("gge", CT_Q_GGE, out.gge, {});
    make("general_ked", CT_Q_GENERAL_KED, out.general_ked, {});
    make("ked_alpha", CT_Q_KED_ALPHA, out.ked_alpha, {});
    ct_eval_params par{};
    par.gge_a = a;
    par.gge_b = b;
    par.ked_a = alpha;
    par.ked_alpha = alpha;
    par.mixed_tol = mixed_tol;
    if (route == "auto")
      par.route = CT_ROUTE_AUTO;
    else if (route == "density")
      par.route = CT_ROUTE_DENSITY;
    else if (route == "factor")
      par.route = CT_ROUTE_FACTOR;
    else
      throw std::invalid_argument("route must be 'auto', 'density' or 'factor'");
    ct_eval_report rep{};
    const int prec = parse_precision(precision);
    int st;
    {
      py::gil_scoped_release nogil;
      if (cube)
        st = ct_eval_cubic(h_, cube->origin, cube->axes, cube->shape, q, prec, &par, &out, &rep);
      else
        st = ct_eval(h_, pts, n, CT_LAYOUT_ROW, q, prec, &par, &out, &rep);
    }
    check(st);
    last_report_ = rep;
    return res;
  }

  ct_mol* h_ = nullptr;
  std::string path_;
  std::vector<int32_t> numbers_override_;
  ct_eval_report last_report_{};
};

py::array_t<double> norm_rows(const Points& v) {
  const int64_t n = n_points(v);
  py::array_t<double> out(n);
  int st;
  {
    py::gil_scoped_release nogil;
    st = ct_norm_rows(v.data(), n, out.mutable_data());
  }
  check(st);
  return out;
}

}  // namespace

PYBIND11_MODULE(_chemtools_amd, m) {
  m.doc() = "MI355X evaluator of Gaussian-basis electron-density descriptors (HIP/CDNA4).";
  const char* P64 = "fp64";
  py::class_<Molecule>(m, "Molecule")
      .def(py::init<const std::string&>(), py::arg("file_path"),
           "Load a wavefunction from a Gaussian formatted checkpoint (.fchk) file.")
      .def(py::init<py::array_t<int32_t, py::array::c_style | py::array::forcecast>,
                    py::array_t<int32_t, py::array::c_style | py::array::forcecast>,
                    py::array_t<int32_t, py::array::c_style | py::array::forcecast>,
                    py::array_t<int32_t, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>, py::object>(),
           py::arg("shell_atom"), py::arg("shell_l"), py::arg("shell_nprim"), py::arg("shell_prim_off"),
           py::arg("exps"), py::arg("coefs"), py::arg("atom_xyz"), py::arg("charges"), py::arg("P"),
           py::arg("numbers") = py::none(), "Build from an in-memory shell table and density matrix.")
      .def("compute_density", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_RHO, {}, pr); },
           py::arg("points"), py::arg("precision") = P64, "Electron density rho(r); returns (N,).")
      .def("compute_electron_density", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_RHO, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_gradient", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_GRAD, {3}, pr); },
           py::arg("points"), py::arg("precision") = P64, "Gradient of rho; returns (N, 3).")
      .def("compute_electron_density_gradient", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_GRAD, {3}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_hessian", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_HESS, {3, 3}, pr); },
           py::arg("points"), py::arg("precision") = P64, "Hessian of rho; returns (N, 3, 3).")
      .def("compute_electron_density_hessian", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_HESS, {3, 3}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_laplacian", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_LAP, {}, pr); },
           py::arg("points"), py::arg("precision") = P64, "Laplacian of rho (trace of the Hessian); returns (N,).")
      .def("compute_laplacian_electron_density", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_LAP, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_positive_definite_kinetic_energy_density",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_TAU_PD, {}, pr); },
           py::arg("points"), py::arg("precision") = P64, "tau_PD = 1/2 sum_k d_k phi P d_k phi.")
      .def("compute_general_kinetic_energy_density",
           [](Molecule& s, const Points& p, double alpha, const std::string& pr) { return s.one(p, CT_Q_GENERAL_KED, {}, pr, 0, 0, alpha); },
           py::arg("points"), py::arg("alpha"), py::arg("precision") = P64,
           "Bound, as in the reference, to the general(ish) KED: tau_PD + (alpha - 1)/4 lap.")
      .def("compute_general_kinetic_energy_density_alpha",
           [](Molecule& s, const Points& p, double alpha, const std::string& pr) { return s.one(p, CT_Q_KED_ALPHA, {}, pr, 0, 0, alpha); },
           py::arg("points"), py::arg("alpha"), py::arg("precision") = P64,
           "t_alpha = tau_PD + alpha lap (the reference's C++ evaluate_general_kinetic_energy_density).")
      .def("compute_norm_of_vector", [](Molecule&, const Points& v) { return norm_rows(v); }, py::arg("array"))
      .def("compute_reduced_density_gradient",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_RDG, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_weizsacker_ked", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_WEIZSACKER, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_thomas_fermi_ked", [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_THOMAS_FERMI, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_general_gradient_expansion_ked",
           [](Molecule& s, const Points& p, double a, double b, const std::string& pr) { return s.one(p, CT_Q_GGE, {}, pr, a, b); },
           py::arg("points"), py::arg("a"), py::arg("b"), py::arg("precision") = P64)
      .def("compute_gradient_expansion_ked",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_GGE, {}, pr, 1.0 / 9.0, 1.0 / 6.0); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_empirical_gradient_expansion_ked",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_GGE, {}, pr, 1.0 / 5.0, 1.0 / 6.0); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_general_ked",
           [](Molecule& s, const Points& p, double a, const std::string& pr) { return s.one(p, CT_Q_GENERAL_KED, {}, pr, 0, 0, a); },
           py::arg("points"), py::arg("a"), py::arg("precision") = P64)
      .def("compute_schrodinger_ked",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_GENERAL_KED, {}, pr, 0, 0, 0.0); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_shannon_information_density",
           [](Molecule& s, const Points& p, const std::string& pr) { return s.one(p, CT_Q_SHANNON, {}, pr); },
           py::arg("points"), py::arg("precision") = P64)
      .def("compute_electrostatic_potential", &Molecule::esp, py::arg("points"), py::arg("precision") = P64)
      .def("compute", &Molecule::compute, py::arg("points"), py::arg("quantities"), py::arg("precision") = P64,
           py::arg("a") = 0.0, py::arg("b") = 0.0, py::arg("alpha") = 0.0, py::arg("mixed_tol") = 0.0,
           py::arg("route") = "auto",
           "Evaluate several quantities in one fused pass; returns a dict.  route (mixed precision): 'auto', "
           "'density' (contract with P) or 'factor' (contract with the pivoted Cholesky factor of P).")
      .def("compute_on_cubic_grid", &Molecule::compute_on_cubic_grid, py::arg("origin"), py::arg("axes"),
           py::arg("shape"), py::arg("quantities"), py::arg("precision") = P64, py::arg("a") = 0.0,
           py::arg("b") = 0.0, py::arg("alpha") = 0.0, py::arg("mixed_tol") = 0.0, py::arg("route") = "auto",
           "Several quantities on the uniform grid origin + i axes[0] + j axes[1] + k axes[2], generated on the "
           "GPU; arrays of shape (n0, n1, n2, ...).")
      .def("compute_electron_density_on_cubic_grid", &Molecule::density_cubic, py::arg("origin"), py::arg("axes"),
           py::arg("shape"), py::arg("disp") = false, py::arg("precision") = P64,
           "Electron density on a uniform grid, shape (n0, n1, n2) (the reference's test spelling).")
      .def("compute_electron_density_cubic", &Molecule::density_cubic, py::arg("origin"), py::arg("axes"),
           py::arg("shape"), py::arg("disp") = false, py::arg("precision") = P64,
           "Electron density on a uniform grid, shape (n0, n1, n2) (Molecule::compute_electron_density_cubic).")
      .def("backward", &Molecule::backward, py::arg("points"), py::arg("dL_drho"), py::arg("precision") = P64,
           "Adjoint of rho: returns (dL/dP, dL/dR).")
      .def("set_coordinates", &Molecule::set_coordinates)
      .def("set_density_matrix", &Molecule::set_density_matrix)
      .def("set_workspace_limit", &Molecule::set_workspace_limit)
      .def("set_screening", &Molecule::set_screening, py::arg("eps"),
           "Spatial screening of the factor route: basis-function blocks whose values and derivatives "
           "are at most eps on a point tile are skipped (0 turns it off; default 1e-15; FP64 at 1e-7 of it).")
      .def("last_report", &Molecule::last_report)
      .def_property_readonly("coordinates", &Molecule::coordinates, "Cartesian coordinates of the atomic centres.")
      .def_property_readonly("numbers", &Molecule::numbers, "Atomic numbers of the atomic centres.")
      .def_property_readonly("charges", &Molecule::charges, "Nuclear (core) charges.")
      .def_property_readonly("density_matrix", &Molecule::density_matrix)
      .def_property_readonly("nbasis", &Molecule::nbasis)
      .def_property_readonly("handle", &Molecule::handle)
      .def("get_file_path", &Molecule::file_path, "Get the wavefunction file path.");
}

// grid_search -- bandwidth selection by cross-validation over a grid, then
// evaluation of the chosen (and the adaptive) density over a plotting grid.
//
// Same configuration keys, input format and output files as the reference's
// tool (grid_search/grid_search.cc:40-96, cv_utils_impl.h, cv_utils.cpp):
//   usage: grid_search [--key=value ...] config_fname
// The configuration file holds `key = value` lines ('#' starts a comment);
// a `--key=value` (or `--key value`) on the command line takes precedence.
// Steps (grid_search.cc:158-552):
//   1. read "x y w" rows, write the points for a scatter plot;
//   2. KernelDensity<2, EpanechnikovProductKernel2d<float>, double> with the
//      base bandwidths, made adaptive first when use_adaptive_cv is set;
//   3. unless skip_cross_validation: least-squares CV over the bandwidth grid
//      (automatic: (steps+1)^2 points, ends included; manual: the cross
//      product of two lists).  lsq_conv scores every bandwidth pair in one
//      batched GPU pipeline (lsq_convolution_cross_validate_grid) instead of
//      one dual-tree pass per pair; lsq_numint (forced for adaptive CV) runs
//      per pair.  The smallest score wins (first one on ties);
//   4. the non-adaptive density at the best (or base) bandwidths over the
//      plotting grid;
//   5. the adaptive density (alpha) over the same grid.
// Grid values are written as the reference writes them: the x grid, the y
// grid, then the values row by row (y outer, x inner).
#include <CvGrid.h>
#include <KernelDensity.h>
#include <Kernels/EpanechnikovProductKernel2d.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <functional>
#include <iostream>
#include <iterator>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

using KernelType = bbrcit::EpanechnikovProductKernel2d<float>;
using KernelDensityType = bbrcit::KernelDensity<2, KernelType, double>;
using DataPointType = KernelDensityType::DataPointType;
using Clock = std::chrono::high_resolution_clock;

// ---- options ------------------------------------------------------------------
class Options {
 public:
  void set_if_absent(const std::string &k, const std::string &v) { values_.emplace(k, v); }
  bool has(const std::string &k) const { return values_.count(k) != 0; }
  const std::string &str(const std::string &k) const {
    auto it = values_.find(k);
    if (it == values_.end()) throw std::invalid_argument("missing option '" + k + "'");
    return it->second;
  }
  double dbl(const std::string &k) const { return parse<double>(k, [](const std::string &s, size_t *p) {
                                             return std::stod(s, p); }); }
  int integer(const std::string &k) const { return parse<int>(k, [](const std::string &s, size_t *p) {
                                              return std::stoi(s, p); }); }
  bool boolean(const std::string &k) const {
    const std::string v = str(k);
    if (v == "true" || v == "1" || v == "yes" || v == "on") return true;
    if (v == "false" || v == "0" || v == "no" || v == "off") return false;
    throw std::invalid_argument("option '" + k + "': not a boolean: " + v);
  }

 private:
  template <class T, class F>
  T parse(const std::string &k, F f) const {
    const std::string v = str(k);
    size_t pos = 0;
    T r;
    try {
      r = f(v, &pos);
    } catch (const std::exception &) {
      throw std::invalid_argument("option '" + k + "': cannot convert '" + v + "'");
    }
    if (pos != v.size()) throw std::invalid_argument("option '" + k + "': cannot convert '" + v + "'");
    return r;
  }
  std::map<std::string, std::string> values_;
};

std::string trim(const std::string &s) {
  const auto b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  const auto e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

void read_config(std::istream &in, Options &opt) {
  std::string line;
  while (std::getline(in, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty() || line[0] == '[') continue;
    const auto eq = line.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("configuration line without '=': " + line);
    opt.set_if_absent(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
  }
}

// Split on single spaces; empty tokens fail to convert, as in the
// reference's tokenize<double> (custom_program_option_utils.cpp:11-43).
std::vector<double> tokenize_doubles(const std::string &line) {
  std::vector<double> out;
  size_t b = 0;
  while (true) {
    const size_t e = line.find(' ', b);
    const std::string tok = line.substr(b, e == std::string::npos ? std::string::npos : e - b);
    size_t pos = 0;
    const double v = std::stod(tok, &pos);
    if (pos != tok.size()) throw std::invalid_argument("cannot convert '" + tok + "' to a bandwidth");
    out.push_back(v);
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return out;
}

// ---- data and grids (cv_utils_impl.h:27-59, cv_utils.cpp:6-22) ------------------
std::vector<DataPointType> read_2dpoints(const std::string &fname) {
  std::ifstream fin(fname);
  if (!fin) throw std::ios_base::failure("cannot open file " + fname + ". ");
  std::vector<DataPointType> data;
  double x1 = 0.0, x2 = 0.0, w = 1.0;
  while (fin >> x1 >> x2 >> w) data.push_back(DataPointType({x1, x2}, KernelDensityType::AttributeType(w)));
  return data;
}

std::vector<DataPointType> generate_2dgrid(double start_x, double end_x, int steps_x, double start_y, double end_y,
                                           int steps_y) {
  std::vector<DataPointType> grid;
  const double dx = (end_x - start_x) / steps_x, dy = (end_y - start_y) / steps_y;
  for (int j = 0; j <= steps_y; ++j)
    for (int i = 0; i <= steps_x; ++i) grid.push_back(DataPointType({start_x + i * dx, start_y + j * dy}));
  return grid;
}

void write_2dgrid_values(std::ostream &os, const std::vector<double> &values, double start_x, double end_x,
                         int steps_x, double start_y, double end_y, int steps_y) {
  const double dx = (end_x - start_x) / steps_x, dy = (end_y - start_y) / steps_y;
  for (int i = 0; i <= steps_x; ++i) os << start_x + i * dx << " ";
  os << std::endl;
  for (int i = 0; i <= steps_y; ++i) os << start_y + i * dy << " ";
  os << std::endl;
  for (size_t i = 0; i < values.size(); ++i) {
    if (i % (steps_x + 1) == 0 && i) os << std::endl;
    os << values[i] << " ";
  }
  os << std::endl;
}

// evaluated queries come back in query-tree order: order them y first, then
// x (cv_utils_impl.h:69-89), and write their values
void write_2dgrid_points(std::ostream &os, std::vector<DataPointType> &pts, double start_x, double end_x,
                         int steps_x, double start_y, double end_y, int steps_y) {
  std::sort(pts.begin(), pts.end(), [](const DataPointType &a, const DataPointType &b) {
    if (a[1] != b[1]) return a[1] < b[1];
    return a[0] < b[0];
  });
  std::vector<double> values(pts.size());
  for (size_t i = 0; i < pts.size(); ++i) values[i] = pts[i].attributes().value();
  write_2dgrid_values(os, values, start_x, end_x, steps_x, start_y, end_y, steps_y);
}

void write_to(const std::string &fname, const std::function<void(std::ostream &)> &fn) {
  std::ofstream fout(fname);
  if (!fout) throw std::ios_base::failure("cannot write file " + fname + ". ");
  fn(fout);
  std::cout << "  => wrote results to: " << fname << ". \n" << std::endl;
}

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// ---- the search ------------------------------------------------------------------
void grid_search(const Options &opt) {
  const double rel_tol = opt.dbl("rel_tol"), abs_tol = opt.dbl("abs_tol");
  std::cout << "+ performance parameters: \n\n";
  std::cout << "  relative tolerance: " << rel_tol << "\n  absolute tolerance: " << abs_tol << "\n" << std::endl;
  const int refpt_max_leaf_size = opt.integer("refpt_max_leaf_size");
  const int qgrid_max_leaf_size = opt.integer("qgrid_max_leaf_size");
  std::cout << "  reference point tree max leaf size: " << refpt_max_leaf_size << "\n";
  std::cout << "  query grid tree max leaf size: " << qgrid_max_leaf_size << "\n" << std::endl;

  // 1. reference data
  std::cout << "+ reading in reference data. \n" << std::endl;
  std::vector<DataPointType> data = read_2dpoints(opt.str("input_refpts_fname"));
  std::cout << "  => found " << data.size() << " points. \n" << std::endl;
  write_to(opt.str("output_scatter_fname"), [&](std::ostream &os) {
    for (const auto &p : data) os << p[0] << " " << p[1] << std::endl;
  });

  // 2. the estimator
  const bool use_adaptive_cv = opt.boolean("use_adaptive_cv");
  const bool skip_cross_validation = opt.boolean("skip_cross_validation");
  const double base_bwx = opt.dbl("base_bwx"), base_bwy = opt.dbl("base_bwy");
  const double alpha = opt.dbl("alpha");
  std::cout << "+ constructing kernel density estimator. \n\n";
  std::cout << "  base bandwidth x: " << base_bwx << "\n  base bandwidth y: " << base_bwy << "\n" << std::endl;
  auto t0 = Clock::now();
  KernelDensityType kde(data, refpt_max_leaf_size);
  std::cout << "  => running time: " << ms_since(t0) << " ms. \n" << std::endl;
  kde.kernel().set_bandwidths(base_bwx, base_bwy);
  std::cout << "  perform adaptive density cross validation: " << (use_adaptive_cv ? "true" : "false") << "\n";
  std::cout << "  skip cross validation: " << (skip_cross_validation ? "true" : "false") << "\n" << std::endl;
  if (use_adaptive_cv && !skip_cross_validation) {
    std::cout << "  converting to adaptive density for cross validation (alpha " << alpha << "). \n" << std::endl;
    t0 = Clock::now();
    kde.adapt_density(alpha, rel_tol, abs_tol);
    std::cout << "  => running time: " << ms_since(t0) << " ms. \n" << std::endl;
  }

  // 3. grid search
  const bool use_manual_bwgrid = opt.boolean("use_manual_bwgrid");
  std::string cv_method = opt.str("cv_method");
  const double start_bwx = opt.dbl("start_bwx"), end_bwx = opt.dbl("end_bwx");
  const double start_bwy = opt.dbl("start_bwy"), end_bwy = opt.dbl("end_bwy");
  const int steps_bwx = opt.integer("steps_bwx"), steps_bwy = opt.integer("steps_bwy");
  const double start_qix = opt.dbl("start_qix"), end_qix = opt.dbl("end_qix");
  const double start_qiy = opt.dbl("start_qiy"), end_qiy = opt.dbl("end_qiy");
  const int steps_qix = opt.integer("steps_qix"), steps_qiy = opt.integer("steps_qiy");
  double best_bwx = base_bwx, best_bwy = base_bwy;
  if (!skip_cross_validation) {
    if (!use_adaptive_cv) {
      if (cv_method != "lsq_conv" && cv_method != "lsq_numint")
        throw std::invalid_argument("grid_search: unrecognized cv method. ");
    } else {
      cv_method = "lsq_numint";
    }
    if (cv_method == "lsq_conv") {
      std::cout << "+ using convolution based least squares cross validation. \n" << std::endl;
    } else {
      std::cout << "+ using numerical integration based least squares cross validation. \n";
      std::cout << "    grid dimensions: " << steps_qix << " by " << steps_qiy << "\n" << std::endl;
    }
    std::vector<std::pair<double, double>> bandwidth_grid;
    if (!use_manual_bwgrid) {
      bandwidth_grid = bbrcit::generate_2d_bandwidth_grid(start_bwx, end_bwx, steps_bwx, start_bwy, end_bwy,
                                                          steps_bwy);
      std::cout << "  using automatic bandwidth grid: " << steps_bwx << " by " << steps_bwy << "\n" << std::endl;
    } else {
      bandwidth_grid = bbrcit::make_crossed_vector_pairs(tokenize_doubles(opt.str("manual_bwx")),
                                                         tokenize_doubles(opt.str("manual_bwy")));
      std::cout << "  using manual bandwidth grid: " << opt.str("manual_bwx") << " x " << opt.str("manual_bwy")
                << "\n" << std::endl;
    }
    std::cout << "  grid searching: \n" << std::endl;
    t0 = Clock::now();
    std::vector<double> cv_scores;
    if (cv_method == "lsq_conv") {
      cv_scores = bbrcit::lsq_convolution_cross_validate_grid(kde, bandwidth_grid);
    } else {
      for (const auto &p : bandwidth_grid) {
        kde.kernel().set_bandwidths(p.first, p.second);
        cv_scores.push_back(bbrcit::lsq_numint_cross_validate(kde, start_qix, end_qix, steps_qix, start_qiy, end_qiy,
                                                              steps_qiy, rel_tol, abs_tol, qgrid_max_leaf_size));
      }
    }
    const double elapsed = ms_since(t0);
    double best_cv = std::numeric_limits<double>::max();
    for (size_t k = 0; k < bandwidth_grid.size(); ++k) {
      if (cv_scores[k] < best_cv) {
        best_cv = cv_scores[k];
        best_bwx = bandwidth_grid[k].first;
        best_bwy = bandwidth_grid[k].second;
      }
      std::cout << "  (" << bandwidth_grid[k].first << ", " << bandwidth_grid[k].second << ") " << cv_scores[k]
                << std::endl;
    }
    std::cout << "\n  => best bandwidth: (" << best_bwx << ", " << best_bwy << "), cv score = " << best_cv << "\n"
              << std::endl;
    std::cout << "  => running time per iteration: " << elapsed / bandwidth_grid.size() << " ms  \n" << std::endl;
    if (!use_manual_bwgrid)
      write_to(opt.str("output_gridsearch_fname"), [&](std::ostream &os) {
        write_2dgrid_values(os, cv_scores, start_bwx, end_bwx, steps_bwx, start_bwy, end_bwy, steps_bwy);
      });
  }

  // 4. the cross-validated, non-adaptive density over the plotting grid
  const bool use_gridsearch_best = opt.boolean("use_gridsearch_best") && !skip_cross_validation;
  const double start_qx = opt.dbl("start_qx"), end_qx = opt.dbl("end_qx");
  const double start_qy = opt.dbl("start_qy"), end_qy = opt.dbl("end_qy");
  const int steps_qx = opt.integer("steps_qx"), steps_qy = opt.integer("steps_qy");
  kde.unadapt_density();
  if (use_gridsearch_best && !use_adaptive_cv)
    kde.kernel().set_bandwidths(best_bwx, best_bwy);
  else
    kde.kernel().set_bandwidths(base_bwx, base_bwy);
  const std::vector<DataPointType> grid = generate_2dgrid(start_qx, end_qx, steps_qx, start_qy, end_qy, steps_qy);
  std::cout << "+ evaluating cross validated non-adapted kernel density over plotting grid: " << steps_qx << "x"
            << steps_qy << "\n" << std::endl;
  std::vector<DataPointType> queries = grid;
  t0 = Clock::now();
  kde.eval(queries, rel_tol, abs_tol, qgrid_max_leaf_size);
  std::cout << "  => running time: " << ms_since(t0) << " ms. \n" << std::endl;
  write_to(opt.str("output_eval_fname"), [&](std::ostream &os) {
    write_2dgrid_points(os, queries, start_qx, end_qx, steps_qx, start_qy, end_qy, steps_qy);
  });

  // 5. the adaptive density over the plotting grid
  double before_bwx = base_bwx, before_bwy = base_bwy;
  if (!use_adaptive_cv && use_gridsearch_best) {
    before_bwx = best_bwx;
    before_bwy = best_bwy;
  }
  kde.kernel().set_bandwidths(before_bwx, before_bwy);
  std::cout << "+ converting to adaptive density for evaluation (" << before_bwx << ", " << before_bwy
            << "), alpha " << alpha << ". \n" << std::endl;
  t0 = Clock::now();
  kde.adapt_density(alpha, rel_tol, abs_tol);
  std::cout << "  => running time: " << ms_since(t0) << " ms. \n" << std::endl;
  double after_bwx = before_bwx, after_bwy = before_bwy;
  if (use_adaptive_cv) {
    after_bwx = use_gridsearch_best ? best_bwx : opt.dbl("adapted_base_bwx");
    after_bwy = use_gridsearch_best ? best_bwy : opt.dbl("adapted_base_bwy");
  }
  kde.kernel().set_bandwidths(after_bwx, after_bwy);
  std::cout << "+ evaluating cross validated adapted kernel density over plotting grid (" << after_bwx << ", "
            << after_bwy << "). \n" << std::endl;
  queries = grid;
  t0 = Clock::now();
  kde.eval(queries, rel_tol, abs_tol, qgrid_max_leaf_size);
  std::cout << "  => running time: " << ms_since(t0) << " ms. \n" << std::endl;
  write_to(opt.str("output_adaptive_eval_fname"), [&](std::ostream &os) {
    write_2dgrid_points(os, queries, start_qx, end_qx, steps_qx, start_qy, end_qy, steps_qy);
  });
}

}  // namespace

int main(int argc, char **argv) {
  try {
    Options opt;
    std::string config_fname;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a == "--help" || a == "-h") {
        config_fname.clear();
        break;
      }
      if (a.rfind("--", 0) == 0) {
        a = a.substr(2);
        const auto eq = a.find('=');
        if (eq != std::string::npos) {
          opt.set_if_absent(a.substr(0, eq), a.substr(eq + 1));
        } else if (i + 1 < argc) {
          opt.set_if_absent(a, argv[++i]);
        } else {
          throw std::invalid_argument("option --" + a + " needs a value");
        }
      } else {
        config_fname = a;
      }
    }
    if (config_fname.empty()) {
      std::cout << "\nUsage: grid_search [--key=value ...] config_fname\n" << std::endl;
      return 0;
    }
    std::ifstream fin(config_fname);
    if (!fin) {
      std::cout << "cannot open config file: " << config_fname << std::endl;
      return 0;
    }
    read_config(fin, opt);
    grid_search(opt);
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  } catch (...) {
    std::cerr << "Exception of unknown type!\n";
    return 1;
  }
  return 0;
}

// Drop-in replacement of the reference's single-device experiment driver.
//
// Same command line (src/simulator/single_dev_expt/simulator_expt_platform.cu:312-354,
// src/single_dev_expt_main.cpp:5-17), same init-folder inputs (flag/rho/vx/vy/vz.dat),
// same per-window stdout line `speed = %.4f (MLUPS)` (:654-659) and the same dump
// files `<dumpFolder>/<field>_<step>.dat` (:572-606).  Everything below the
// command line goes through the C ABI of include/culbm.h.
//
// Options added by this engine: --storage {fp32,fp16,bf16}, --storageScale S,
// --waveDepth D, --initDDF {rest,feq}, --rounding {nearest,stochastic}, --json.
#include <culbm.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Options {
  culbm_config cfg{};
  int32_t dstep = 100;
  int32_t nstep = 1000;
  int32_t vel_set = 0;
  std::string init_folder;
  std::string dump_folder = "data/left_inlet_right_outlet_256_256_256_output";
  bool dump_rho = false, dump_vx = false, dump_vy = false, dump_vz = false;
  int32_t init_ddf = CULBM_INIT_REST;
  bool json = false;
  bool dry_run = false;
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

const char* kUsage =
    "Single Device Expt Fluid Solver (MI355X)\n"
    "Usage: single_dev_expt_main [OPTIONS]\n\n"
    "Options:\n"
    "  -h,--help                 Print this help message and exit\n"
    "  --devId INT               Device index of the GPU [0]\n"
    "  --dstep UINT              Delta time steps of the simulation. [100]\n"
    "  --nstep UINT              Total time steps of the simulation. [1000]\n"
    "  --invTau FLOAT            Reciprocal of Tau. [0.5]\n"
    "  --blockDim [x,y,z]        Dimension of blocks per grid. [[32,16,2]]\n"
    "  --gridDim [x,y,z]         Dimension of grids per kernel. [[2,3,19]]\n"
    "  --domainDim [x,y,z]       Dimension of the domain. [[256,256,256]]\n"
    "  --streamPolicy INT        {[0]:PULL_STREAM, [1]:INPLACE_STREAM} [0]\n"
    "  --optPolicy INT           {[0]:NONE, [1]:HALO_BLOCKING_STATIC_L2, [2]:HALO_BLOCKING_L1L2,\n"
    "                            [3]:HALO_BLOCKING_DYNAMIC_L2} [0]\n"
    "  --velSet INT              The velocity set of the solver. {[0]:D3Q27} [0]\n"
    "  --initStateFolder TEXT    The folder contains initial state of the simulation.\n"
    "  --dumpFolder TEXT         The folder contains final output.\n"
    "  --dumpRho/--dumpVx/--dumpVy/--dumpVz   Dump the field every `dstep` steps.\n"
    "  --useRingDDFBuf           One global DDF buffer for Dynamic L2 in-place streaming.\n"
    "  --innerLoop INT           # of iterations in halo-blocking. [5]\n"
    "  --storage TEXT            Population storage {fp32,fp16,bf16} [fp32]\n"
    "  --storageScale FLOAT      16-bit storage scale (power of two, 0 = per run) [0]\n"
    "  --waveDepth INT           Planes per temporal-blocking chunk (0 = automatic) [0]\n"
    "  --initDDF TEXT            Populations at start {rest,feq} [rest]\n"
    "  --rounding TEXT           16-bit stores {nearest,stochastic} [nearest]\n"
    "  --json                    Also print one JSON line per window\n"
    "  --dryRun                  Validate, print the parsed configuration as JSON and exit\n";

long parse_int(const std::string& opt, const std::string& v) {
  char* end = nullptr;
  errno = 0;
  const long x = std::strtol(v.c_str(), &end, 10);
  if (errno || end == v.c_str() || *end != '\0')
    throw ParseError(opt + ": Value " + v + " could not be converted");
  return x;
}

double parse_float(const std::string& opt, const std::string& v) {
  char* end = nullptr;
  errno = 0;
  const double x = std::strtod(v.c_str(), &end);
  if (errno || end == v.c_str() || *end != '\0')
    throw ParseError(opt + ": Value " + v + " could not be converted");
  return x;
}

// Vectors as `[x,y,z]`, `x,y,z` or three separate tokens.
void parse_vec3(const std::string& opt, std::vector<std::string>& args, size_t& k,
                const std::string* inline_val, int32_t out[3]) {
  std::string text;
  if (inline_val) {
    text = *inline_val;
  } else {
    if (k + 1 >= args.size()) throw ParseError(opt + " requires 3 values");
    text = args[++k];
  }
  std::string s;
  for (char ch : text)
    if (ch != '[' && ch != ']' && ch != ' ') s += ch;
  std::vector<std::string> parts;
  size_t pos = 0;
  while (true) {
    const size_t c = s.find(',', pos);
    parts.push_back(s.substr(pos, c == std::string::npos ? std::string::npos : c - pos));
    if (c == std::string::npos) break;
    pos = c + 1;
  }
  if (parts.size() == 1 && !inline_val) {
    // separate tokens: x y z
    if (k + 2 >= args.size()) throw ParseError(opt + " requires 3 values");
    parts.push_back(args[++k]);
    parts.push_back(args[++k]);
  }
  if (parts.size() != 3) throw ParseError(opt + " requires 3 values");
  for (int i = 0; i < 3; ++i) {
    const long v = parse_int(opt, parts[i]);
    if (v < 0) throw ParseError(opt + ": Value " + parts[i] + " could not be converted");
    out[i] = int32_t(v);
  }
}

Options parse(int argc, char** argv) {
  Options o;
  culbm_config_default(&o.cfg);
  std::vector<std::string> args(argv + 1, argv + argc);
  for (size_t k = 0; k < args.size(); ++k) {
    std::string opt = args[k];
    std::string inl;
    bool has_inline = false;
    const size_t eq = opt.find('=');
    if (opt.rfind("--", 0) == 0 && eq != std::string::npos) {
      inl = opt.substr(eq + 1);
      opt = opt.substr(0, eq);
      has_inline = true;
    }
    auto value = [&]() -> std::string {
      if (has_inline) return inl;
      if (k + 1 >= args.size()) throw ParseError(opt + " requires an argument");
      return args[++k];
    };
    auto flag = [&]() -> bool {
      if (!has_inline) return true;
      if (inl == "true" || inl == "1") return true;
      if (inl == "false" || inl == "0") return false;
      throw ParseError(opt + ": flag value " + inl + " not understood");
    };
    if (opt == "-h" || opt == "--help") {
      std::cout << kUsage;
      std::exit(0);
    } else if (opt == "--devId") {
      o.cfg.device = int32_t(parse_int(opt, value()));
    } else if (opt == "--dstep") {
      const long v = parse_int(opt, value());
      if (v < 0) throw ParseError(opt + ": Value must be non-negative");
      o.dstep = int32_t(v);
    } else if (opt == "--nstep") {
      const long v = parse_int(opt, value());
      if (v < 0) throw ParseError(opt + ": Value must be non-negative");
      o.nstep = int32_t(v);
    } else if (opt == "--invTau") {
      o.cfg.inv_tau = float(parse_float(opt, value()));
    } else if (opt == "--blockDim") {
      parse_vec3(opt, args, k, has_inline ? &inl : nullptr, o.cfg.block_dim);
    } else if (opt == "--gridDim") {
      parse_vec3(opt, args, k, has_inline ? &inl : nullptr, o.cfg.grid_dim);
    } else if (opt == "--domainDim") {
      int32_t d[3];
      parse_vec3(opt, args, k, has_inline ? &inl : nullptr, d);
      o.cfg.nx = d[0];
      o.cfg.ny = d[1];
      o.cfg.nz = d[2];
    } else if (opt == "--streamPolicy") {
      o.cfg.stream_policy = int32_t(parse_int(opt, value()));
    } else if (opt == "--optPolicy") {
      o.cfg.opt_policy = int32_t(parse_int(opt, value()));
    } else if (opt == "--velSet") {
      o.vel_set = int32_t(parse_int(opt, value()));
      if (o.vel_set != 0) throw ParseError("--velSet: only 0 (D3Q27) is available");
    } else if (opt == "--initStateFolder") {
      o.init_folder = value();
    } else if (opt == "--dumpFolder") {
      o.dump_folder = value();
    } else if (opt == "--dumpRho") {
      o.dump_rho = flag();
    } else if (opt == "--dumpVx") {
      o.dump_vx = flag();
    } else if (opt == "--dumpVy") {
      o.dump_vy = flag();
    } else if (opt == "--dumpVz") {
      o.dump_vz = flag();
    } else if (opt == "--useRingDDFBuf") {
      o.cfg.use_ring = flag() ? 1 : 0;
    } else if (opt == "--innerLoop") {
      o.cfg.inner_loop = int32_t(parse_int(opt, value()));
    } else if (opt == "--storage") {
      const std::string v = value();
      if (v == "fp32") o.cfg.storage = CULBM_STORAGE_FP32;
      else if (v == "fp16") o.cfg.storage = CULBM_STORAGE_FP16;
      else if (v == "bf16") o.cfg.storage = CULBM_STORAGE_BF16;
      else throw ParseError("--storage: expected fp32, fp16 or bf16, got " + v);
    } else if (opt == "--storageScale") {
      o.cfg.storage_scale = float(parse_float(opt, value()));
    } else if (opt == "--waveDepth") {
      o.cfg.wave_depth = int32_t(parse_int(opt, value()));
    } else if (opt == "--initDDF") {
      const std::string v = value();
      if (v == "rest") o.init_ddf = CULBM_INIT_REST;
      else if (v == "feq") o.init_ddf = CULBM_INIT_FEQ;
      else throw ParseError("--initDDF: expected rest or feq, got " + v);
    } else if (opt == "--rounding") {
      const std::string v = value();
      if (v == "nearest") o.cfg.rounding = CULBM_ROUND_NEAREST;
      else if (v == "stochastic") o.cfg.rounding = CULBM_ROUND_STOCHASTIC;
      else throw ParseError("--rounding: expected nearest or stochastic, got " + v);
    } else if (opt == "--json") {
      o.json = flag();
    } else if (opt == "--dryRun") {
      o.dry_run = flag();
    } else {
      throw ParseError("The following argument was not expected: " + args[k]);
    }
  }
  return o;
}

void check(int rc, culbm_ctx* ctx) {
  if (rc != CULBM_OK) throw std::runtime_error(culbm_last_error(ctx));
}

template <class T>
bool read_file(const std::string& path, std::vector<T>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(reinterpret_cast<char*>(out.data()), std::streamsize(out.size() * sizeof(T)));
  return true;
}

class Simulator {
 public:
  explicit Simulator(const Options& o) : o_(o) {
    char msg[512];
    const int v = culbm_validate(&o_.cfg, msg, sizeof(msg));
    if (v != CULBM_OK) throw std::runtime_error(msg);
    // host inputs first (initFlag / initRhoU, simulator_expt_platform.cu:95-118, 200-228),
    // so a missing flag.dat is reported before any device work
    const size_t n = size_t(o_.cfg.nx) * o_.cfg.ny * o_.cfg.nz;
    if (!o_.init_folder.empty()) {
      flag_.resize(n);
      if (!read_file(o_.init_folder + "/flag.dat", flag_))
        throw std::runtime_error("File: " + o_.init_folder + "/flag.dat doesn't exist!");
    }
    const char* names[4] = {"rho", "vx", "vy", "vz"};
    for (int k = 0; k < 4; ++k) {
      fields_[k].resize(n);
      from_file_[k] = !o_.init_folder.empty() &&
                      read_file(o_.init_folder + "/" + names[k] + ".dat", fields_[k]);
      if (!from_file_[k]) fields_[k].clear();
    }
    // the plain pull run covers the domain with blockDim * gridDim exactly
    // (staticPullRun, simulator_expt_platform.cu:962-967), checked after the
    // init files as there; the library picks its own launch shape, so only the
    // drop-in CLI keeps this rule
    if (o_.cfg.stream_policy == CULBM_STREAM_PULL && o_.cfg.opt_policy == CULBM_OPT_NONE) {
      const int32_t dom[3] = {o_.cfg.nx, o_.cfg.ny, o_.cfg.nz};
      for (int k = 0; k < 3; ++k)
        if (int64_t(o_.cfg.block_dim[k]) * o_.cfg.grid_dim[k] != dom[k])
          throw std::runtime_error("Dimension of domain doesn't match computation grid!");
    }
  }

  void init() {
    check(culbm_create(&o_.cfg, &ctx_), nullptr);
    if (!flag_.empty())
      std::cout << "Init flag successfully, from file: " << o_.init_folder << "/flag.dat." << std::endl;
    const char* names[4] = {"rho", "vx", "vy", "vz"};
    const char* dflt[4] = {"1", "0", "0", "0"};
    for (int k = 0; k < 4; ++k) {
      if (from_file_[k])
        std::cout << "Init " << names[k] << " successfully, from file: " << o_.init_folder << "/"
                  << names[k] << ".dat." << std::endl;
      else
        std::cout << "Init " << names[k] << " successfully, default init with value " << dflt[k]
                  << "." << std::endl;
    }
    auto p = [](std::vector<float>& v) { return v.empty() ? nullptr : v.data(); };
    check(culbm_set_state(ctx_, flag_.empty() ? nullptr : flag_.data(), p(fields_[0]), p(fields_[1]),
                          p(fields_[2]), p(fields_[3])),
          ctx_);
    check(culbm_init_ddf(ctx_, o_.init_ddf), ctx_);
    std::cout << (o_.cfg.use_ring ? "Init ring ddf buffer successfully."
                                  : "Init double ddf buffer successfully.")
              << std::endl;
  }

  void print_config() const {
    const culbm_config& c = o_.cfg;
    std::printf(
        "{\"domainDim\": [%d, %d, %d], \"blockDim\": [%d, %d, %d], \"gridDim\": [%d, %d, %d], "
        "\"invTau\": %.9g, \"streamPolicy\": %d, \"optPolicy\": %d, \"innerLoop\": %d, "
        "\"nstep\": %d, \"dstep\": %d, \"devId\": %d, \"useRingDDFBuf\": %s, \"storage\": %d, "
        "\"rounding\": %d, \"initStateFolder\": \"%s\", \"dumpFolder\": \"%s\", \"dump\": [%d, %d, %d, %d]}\n",
        c.nx, c.ny, c.nz, c.block_dim[0], c.block_dim[1], c.block_dim[2], c.grid_dim[0], c.grid_dim[1],
        c.grid_dim[2], double(c.inv_tau), c.stream_policy, c.opt_policy, c.inner_loop, o_.nstep,
        o_.dstep, c.device, c.use_ring ? "true" : "false", c.storage, c.rounding, o_.init_folder.c_str(),
        o_.dump_folder.c_str(), int(o_.dump_rho), int(o_.dump_vx), int(o_.dump_vy), int(o_.dump_vz));
  }
  ~Simulator() { culbm_destroy(ctx_); }

  // The reference's window loop (simulator_expt_platform.cu:963-991): each
  // window's dump is armed as a deferred dump and written by the first step of
  // the next window (no separate pass over the populations), its files once
  // that window has run; the final window's dump runs as its own pass.
  void run() {
    const double cells = double(o_.cfg.nx) * o_.cfg.ny * o_.cfg.nz;
    int64_t step = 0, pending = -1;
    int window = 0;
    while (step < o_.nstep) {
      int32_t done = 0;
      float ms = 0.0f;
      check(culbm_run_window(ctx_, o_.dstep, &done, &ms), ctx_);
      const double mlups = ms > 0.0f ? cells * 1e-6 / (double(ms) / 1000.0) * done : 0.0;
      std::printf("speed = %.4f (MLUPS)\n", mlups);
      if (o_.json)
        std::printf("{\"window\": %d, \"steps\": %d, \"ms\": %.6f, \"mlups\": %.4f}\n", window, done,
                    double(ms), mlups);
      std::fflush(stdout);
      if (pending >= 0) dump_res(pending);  // the window's first step wrote it
      step += done;
      ++window;
      if (step < o_.nstep && done > 0) {
        check(culbm_dump_deferred(ctx_), ctx_);
        pending = step;
      } else {
        check(culbm_dump(ctx_), ctx_);
        dump_res(step);
        pending = -1;
      }
      if (done == 0) break;  // dstep 0 makes no progress, as the reference's loop would spin
    }
  }

 private:
  void dump_res(int64_t step) {
    std::filesystem::create_directories(o_.dump_folder);
    if (!(o_.dump_rho || o_.dump_vx || o_.dump_vy || o_.dump_vz)) return;
    const size_t n = size_t(o_.cfg.nx) * o_.cfg.ny * o_.cfg.nz;
    std::vector<float> f[4];
    const bool on[4] = {o_.dump_rho, o_.dump_vx, o_.dump_vy, o_.dump_vz};
    for (int k = 0; k < 4; ++k)
      if (on[k]) f[k].resize(n);
    auto p = [](std::vector<float>& v) { return v.empty() ? nullptr : v.data(); };
    check(culbm_get_macros(ctx_, p(f[0]), p(f[1]), p(f[2]), p(f[3])), ctx_);
    const char* names[4] = {"rho", "vx", "vy", "vz"};
    for (int k = 0; k < 4; ++k) {
      if (!on[k]) continue;
      const std::string path =
          (std::filesystem::path(o_.dump_folder) / (std::string(names[k]) + "_" + std::to_string(step) + ".dat"))
              .string();
      std::ofstream out(path, std::ios::binary);
      if (!out) throw std::runtime_error("Could not open the file: " + path);
      out.write(reinterpret_cast<const char*>(f[k].data()), std::streamsize(n * sizeof(float)));
    }
  }

  Options o_;
  culbm_ctx* ctx_ = nullptr;
  std::vector<uint32_t> flag_;
  std::vector<float> fields_[4];
  bool from_file_[4] = {false, false, false, false};
};

}  // namespace

int main(int argc, char** argv) {
  Options o;
  try {
    o = parse(argc, argv);
  } catch (const ParseError& e) {
    std::cerr << e.what() << "\nRun with --help for more information." << std::endl;
    return 2;
  }
  try {
    Simulator sim(o);
    if (o.dry_run) {
      sim.print_config();
      return 0;
    }
    sim.init();
    sim.run();
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}

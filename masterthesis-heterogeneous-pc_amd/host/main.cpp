// heterogpc command line (src/main.cpp:17-141): same options, defaults and
// behaviour.  Options are parsed by hand (the reference uses
// boost::program_options): long options take `--name value` or
// `--name=value` and may be abbreviated to any unambiguous prefix; short
// options take `-x value` or `-xvalue`.
#include <stdlib.h>

#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "skeleton.hpp"

using namespace std;

const string DEFAULT_INPUT_FILE = "../../data/cooling_house.csv";

namespace {

struct Opt {
  string name;
  char shortName;
  bool takesValue;
  bool multitoken;
  string help;
  string def;
};

const vector<Opt> kOptions = {
    {"help", 0, false, false, "produce a help message", ""},
    {"input-file", 'i', true, false, "input file", DEFAULT_INPUT_FILE},
    {"alpha", 'a', true, false, "alpha value", "0.05"},
    {"observations", 'o', true, false, "observation count", ""},
    {"row-mult", 0, true, false, "maximum rows balanced multiplier level 0,1,4+", "0.2"},
    {"row-mult2", 0, true, false, "maximum rows balanced multiplier level 2", "0.2"},
    {"row-mult3", 0, true, false, "maximum rows balanced multiplier level 3", "0.1"},
    {"max-level", 'm', true, false, "maximum level", "4"},
    {"corr", 0, false, false, "input file is a correlation matrix", ""},
    {"gpus", 'g', true, true, "GPU deviceIds that should be used", ""},
    {"thread-count", 't', true, false, "number of threads used by openMP", ""},
    {"csv-export", 0, true, false, "Export runtimes execution metrics to CSV", ""},
    {"gpu-only", 0, false, false, "execution on gpu only", ""},
    {"cpu-only", 0, false, false, "execution on cpu only", ""},
    {"power9-ats", 0, false, false,
     "Use power9 system allocator for ATS (Address translation service)", ""},
    {"workstealing", 'w', false, false, "use workstealing CPU executor", ""},
    {"print-sepsets", 'p', false, false, "prints-sepsets", ""},
    {"verbose", 'v', false, false, "verbose output", ""},
    {"all-levels", 0, false, false, "also test levels >= 4 without --workstealing", ""},
};

void print_help() {
  cout << "Allowed options:\n";
  for (const auto &o : kOptions) {
    string lhs = "  ";
    lhs += o.shortName ? string("-") + o.shortName + " [ --" + o.name + " ]" : "--" + o.name;
    if (o.takesValue) lhs += o.def.empty() ? " arg" : " arg (=" + o.def + ")";
    if (lhs.size() < 42) lhs.resize(42, ' ');
    cout << lhs << " " << o.help << "\n";
  }
}

const Opt *find_long(const string &name, string &err) {
  const Opt *hit = nullptr;
  int matches = 0;
  for (const auto &o : kOptions) {
    if (o.name == name) return &o;
    if (o.name.compare(0, name.size(), name) == 0) { hit = &o; ++matches; }
  }
  if (matches == 1) return hit;
  err = matches ? "option '--" + name + "' is ambiguous" : "unrecognised option '--" + name + "'";
  return nullptr;
}

const Opt *find_short(char c) {
  for (const auto &o : kOptions)
    if (o.shortName == c) return &o;
  return nullptr;
}

bool is_option(const char *s) {
  return s[0] == '-' && s[1] != '\0' && !(s[1] >= '0' && s[1] <= '9') && s[1] != '.';
}

// name -> values; returns false with err on a parse error
bool parse(int argc, char const *argv[], map<string, vector<string>> &vm, string &err) {
  for (int a = 1; a < argc; ++a) {
    const string arg = argv[a];
    const Opt *o = nullptr;
    string inline_val;
    bool has_inline = false;
    if (arg.rfind("--", 0) == 0) {
      string name = arg.substr(2);
      const size_t eq = name.find('=');
      if (eq != string::npos) { inline_val = name.substr(eq + 1); name = name.substr(0, eq); has_inline = true; }
      o = find_long(name, err);
      if (!o) return false;
    } else if (arg.size() >= 2 && arg[0] == '-') {
      o = find_short(arg[1]);
      if (!o) { err = "unrecognised option '" + arg + "'"; return false; }
      if (arg.size() > 2) { inline_val = arg.substr(2); has_inline = true; }
    } else {
      err = "unexpected positional argument '" + arg + "'";
      return false;
    }
    auto &vals = vm[o->name];
    if (!o->takesValue) {
      if (has_inline) { err = "option '--" + o->name + "' does not take a value"; return false; }
      vals.push_back("");
      continue;
    }
    if (has_inline) {
      vals.push_back(inline_val);
    } else if (a + 1 < argc && !is_option(argv[a + 1])) {
      vals.push_back(argv[++a]);
    } else {
      err = "the required argument for option '--" + o->name + "' is missing";
      return false;
    }
    if (o->multitoken)
      while (a + 1 < argc && !is_option(argv[a + 1])) vals.push_back(argv[++a]);
  }
  return true;
}

string value(map<string, vector<string>> &vm, const string &name) {
  auto it = vm.find(name);
  if (it != vm.end() && !it->second.empty()) return it->second.back();
  for (const auto &o : kOptions)
    if (o.name == name) return o.def;
  return "";
}

}  // namespace

int main(int argc, char const *argv[]) {
  map<string, vector<string>> vm;
  string err;
  if (!parse(argc, argv, vm, err)) {
    cerr << "Error: " << err << "\n";
    return 1;
  }
  if (vm.count("help")) {
    print_help();
    return 0;
  }
  const bool verbose = vm.count("verbose") != 0;
  int numberOfThreads = defaultThreadCount();
  try {
    if (vm.count("thread-count")) numberOfThreads = stoi(value(vm, "thread-count"));
    const string inputFile = value(vm, "input-file");
    const double alpha = stod(value(vm, "alpha"));
    const int maxLevel = stoi(value(vm, "max-level"));
    const bool workstealing = vm.count("workstealing") != 0;
    const string csvExportFile = vm.count("csv-export") ? value(vm, "csv-export") : "";
    SkeletonCalculator skeletonCalculator(maxLevel, alpha, workstealing, csvExportFile,
                                          numberOfThreads, verbose);
    if (vm.count("gpu-only")) skeletonCalculator.set_heterogeneity(GPUOnly);
    else if (vm.count("cpu-only")) skeletonCalculator.set_heterogeneity(CPUOnly);
    if (vm.count("gpus")) {
      vector<int> gpus;
      for (auto &g : vm["gpus"]) gpus.push_back(stoi(g));
      skeletonCalculator.set_gpu_list(gpus);
    }
    skeletonCalculator.set_all_levels(vm.count("all-levels") != 0);
    const bool print_sepsets = vm.count("print-sepsets") != 0;
    const bool use_p9_ats = vm.count("power9-ats") != 0;
    if (vm.count("corr")) {
      if (!vm.count("observations")) {
        cout << "Observation count needed with correlation matrix input" << endl;
        return -1;
      }
      skeletonCalculator.add_correlation_matrix(inputFile, stoi(value(vm, "observations")),
                                                use_p9_ats);
    } else {
      skeletonCalculator.add_observations(inputFile, use_p9_ats);
    }
    tuple<float, float, float> thresholds = {stof(value(vm, "row-mult")),
                                             stof(value(vm, "row-mult2")),
                                             stof(value(vm, "row-mult3"))};
    skeletonCalculator.initialize_balancer(thresholds);
    skeletonCalculator.run(print_sepsets);
  } catch (const std::invalid_argument &e) {
    cerr << "Error: the argument is invalid (" << e.what() << ")\n";
    return 1;
  } catch (const std::out_of_range &e) {
    cerr << "Error: the argument is out of range (" << e.what() << ")\n";
    return 1;
  }
  return 0;
}

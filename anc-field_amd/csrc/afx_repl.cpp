// afx-field: interactive command surface over the afx C ABI.
//
// Command-compatible with the reference REPL (paths relative to the
// reference root): anc-field/src/main.cpp:32-143 (commands and their
// argument types) and anc-field/src/commandline.hpp:107-196 (prompt ">> ",
// argument checking and its "Error: ..." messages).  Progress is printed as
// "BEGIN" then "i/steps" every 200 steps, as Simulation::Run does
// (simulation.hpp:403-409, :433), which is what the Python wrapper parses.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <iterator>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/afx.h"

namespace {

int to_int(const std::string& t)
{
    size_t used = 0;
    int v = std::stoi(t, &used);
    if (used < t.size()) throw std::invalid_argument("int");
    return v;
}

float to_float(const std::string& t)
{
    size_t used = 0;
    float v = std::stof(t, &used);
    if (used < t.size()) throw std::invalid_argument("float");
    return v;
}

// "[a,b,c]" -> values (commandline.hpp ToIntList / ToFloatList)
template <typename T>
std::vector<T> to_list(const std::string& t, T (*conv)(const std::string&))
{
    if (t.size() < 2 || t.front() != '[' || t.back() != ']') throw std::invalid_argument("list");
    std::vector<T> out;
    std::stringstream ss(t.substr(1, t.size() - 2));
    std::string item;
    while (std::getline(ss, item, ',')) out.push_back(conv(item));
    return out;
}

// whitespace-separated numbers, rows flattened in order (datafile.hpp:28-66)
bool load_samples(const std::string& path, std::vector<float>& out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string tok;
        while (ss >> tok) out.push_back((float)std::atof(tok.c_str()));
    }
    return true;
}

struct Repl {
    afx_sim* sim = nullptr;
    afx_config cfg{};
    int nmic = 0;
    std::vector<int> devices;

    bool ok(int rc)
    {
        if (rc == AFX_OK) return true;
        std::cout << "Error: " << afx_last_error() << "\n";
        return false;
    }

    void create(int dev)
    {
        if (sim) afx_destroy(sim);
        sim = nullptr;
        afx_config_default(&cfg);
        cfg.device = dev;
        nmic = 0;
        ok(afx_create(&cfg, &sim));
    }

    bool need_sim()
    {
        if (sim) return true;
        std::cout << "Error: no active device (use sd or sdt)\n";
        return false;
    }
};

}  // namespace

int main()
{
    Repl r;
    using Handler = std::function<void(const std::vector<std::string>&)>;
    std::map<std::string, std::pair<Handler, std::vector<std::string>>> cmds;

    cmds["lsd"] = {[&](const std::vector<std::string>&) {
                       int32_t n = 0;
                       afx_device_count(&n);
                       std::cout << "Available compute devices" << std::endl;
                       for (int i = 0; i < n; ++i) {
                           char name[256];
                           if (afx_device_name(i, name, sizeof name) == AFX_OK) std::cout << i << ". " << name << std::endl;
                       }
                   },
                   {}};
    cmds["sd"] = {[&](const std::vector<std::string>& a) {
                      auto ix = to_list<int>(a[0], to_int);
                      if (ix.empty()) return;
                      for (int i : ix) {
                          char name[256];
                          if (!r.ok(afx_device_name(i, name, sizeof name))) return;
                          r.devices.push_back(i);
                          std::cout << name << " added to active devices\n";
                      }
                      r.create(r.devices.front());  // only the first device is used (simulation.hpp:176-178)
                  },
                  {"int_list"}};
    cmds["sdt"] = {[&](const std::vector<std::string>& a) {
                       if (a[0] == "gpu") {
                           char name[256];
                           if (!r.ok(afx_device_name(0, name, sizeof name))) return;
                           r.devices.push_back(0);
                           std::cout << name << " added to active devices\n";
                           r.create(0);
                       } else if (a[0] == "cpu") {
                           std::cout << "Error: this build runs on HIP devices only\n";
                       } else {
                           std::cout << "Valid command arguments are 'cpu' or 'gpu'\n";
                       }
                   },
                   {"string"}};
    cmds["cd"] = {[&](const std::vector<std::string>&) {
                      std::cout << "Compute devices removed\n";
                      r.devices.clear();
                  },
                  {}};
    cmds["load"] = {[&](const std::vector<std::string>& a) {
                        if (r.need_sim()) r.ok(afx_load_kerkythea(r.sim, a[0].c_str()));
                    },
                    {"file"}};
    cmds["addsource"] = {[&](const std::vector<std::string>& a) {
                             if (!r.need_sim()) return;
                             auto p = to_list<float>(a[0], to_float);
                             std::vector<float> s;
                             if (p.size() != 3 || !load_samples(a[1], s)) {
                                 std::cout << "Error: addsource takes [x,y,z] and a readable file\n";
                                 return;
                             }
                             r.ok(afx_add_source(r.sim, p.data(), s.data(), (int64_t)s.size()));
                         },
                         {"float_list", "file"}};
    cmds["addmic"] = {[&](const std::vector<std::string>& a) {
                          if (!r.need_sim()) return;
                          auto p = to_list<float>(a[0], to_float);
                          if (p.size() != 3) {
                              std::cout << "Error: addmic takes [x,y,z]\n";
                              return;
                          }
                          if (r.ok(afx_add_receiver(r.sim, p.data()))) r.nmic++;
                      },
                      {"float_list"}};
    cmds["vis"] = {[&](const std::vector<std::string>&) {
                       std::cout << "Error: live visualisation is not part of this build\n";
                   },
                   {"float"}};
    cmds["run"] = {[&](const std::vector<std::string>& a) {
                       if (!r.need_sim()) return;
                       const int fs = 32000;  // main.cpp:117-118
                       const int steps = (int)(to_float(a[0]) * (float)fs);
                       for (int i = 0; i < steps; i += 200) {
                           const int k = steps - i < 200 ? steps - i : 200;
                           if (!r.ok(afx_run(r.sim, k))) return;
                           if (i == 0) std::cout << "BEGIN" << std::endl;
                           std::cout << i << "/" << steps << std::endl;
                       }
                       std::cout << steps << "/" << steps << std::endl;
                   },
                   {"float"}};
    cmds["save"] = {[&](const std::vector<std::string>& a) {
                        if (!r.need_sim()) return;
                        int64_t n = 0;
                        if (!r.ok(afx_get_receivers(r.sim, nullptr, 0, &n))) return;
                        std::vector<float> buf((size_t)(n > 0 ? n : 1));
                        if (!r.ok(afx_get_receivers(r.sim, buf.data(), n, &n))) return;
                        // one row per receiver, default stream formatting, trailing space,
                        // final blank line (datafile.hpp:83-95)
                        std::ofstream f(a[0]);
                        const int64_t T = r.nmic ? n / r.nmic : 0;
                        for (int m = 0; m < r.nmic; ++m) {
                            for (int64_t t = 0; t < T; ++t) f << buf[(size_t)(m * T + t)] << " ";
                            f << "\n";
                        }
                        f << "\n";
                    },
                    {"string"}};
    cmds["minsize"] = {[&](const std::vector<std::string>& a) {
                           if (!r.need_sim()) return;
                           auto s = to_list<float>(a[0], to_float);
                           if (s.size() != 3) {
                               std::cout << "Error: minsize takes [x,y,z]\n";
                               return;
                           }
                           r.ok(afx_set_min_size(r.sim, s[0], s[1], s[2]));
                       },
                       {"float_list"}};
    cmds["pml"] = {[&](const std::vector<std::string>& a) {
                       if (r.need_sim()) r.ok(afx_set_pml(r.sim, to_int(a[0]), 0, nullptr, 0));
                   },
                   {"int"}};

    bool go = true;
    while (go) {
        std::cout << ">> ";
        std::string line;
        if (!std::getline(std::cin, line)) break;
        std::istringstream iss(line);
        std::vector<std::string> tok{std::istream_iterator<std::string>{iss}, std::istream_iterator<std::string>{}};
        if (tok.empty()) continue;
        if (tok[0] == "quit") break;
        auto it = cmds.find(tok[0]);
        if (it == cmds.end()) {
            std::cout << "Error: unrecognized command\n";
            continue;
        }
        std::vector<std::string> args(tok.begin() + 1, tok.end());
        const auto& types = it->second.second;
        if (args.size() != types.size()) {
            std::cout << "Error: command " << tok[0] << " takes " << types.size() << " argument(s)\n";
            continue;
        }
        bool args_ok = true;
        for (size_t i = 0; i < args.size(); ++i) {
            try {
                if (types[i] == "int") to_int(args[i]);
                else if (types[i] == "float") to_float(args[i]);
                else if (types[i] == "int_list") to_list<int>(args[i], to_int);
                else if (types[i] == "float_list") to_list<float>(args[i], to_float);
                else if (types[i] == "file" && !std::ifstream(args[i])) {
                    std::cout << "Error: file " << args[i] << " doesn't exist\n";
                    args_ok = false;
                }
            } catch (const std::exception&) {
                std::cout << "Error: argument[" << i << "] should be a(n) " << types[i] << "\n";
                args_ok = false;
            }
        }
        if (!args_ok) continue;
        try {
            it->second.first(args);
        } catch (const std::exception& e) {
            std::cout << "Error: " << e.what() << "\n";
        }
        std::cout.flush();
    }
    if (r.sim) afx_destroy(r.sim);
    return EXIT_SUCCESS;
}

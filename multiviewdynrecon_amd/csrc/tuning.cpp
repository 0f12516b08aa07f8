// The launch-shape table (tuning.h).
#include "tuning.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <utility>

namespace mvdr {
namespace {

std::mutex g_mutex;
std::map<std::pair<uint32_t, uint32_t>, LaunchParams> g_table;
uint64_t g_version = 1;
std::once_flag g_env_once;

// Parses the text form and replaces the table (every line checked first).
bool store_table_text(const std::string& text, std::string& err) {
    std::map<std::pair<uint32_t, uint32_t>, LaunchParams> table;
    std::istringstream in(text);
    std::string line;
    int number = 0;
    while (std::getline(in, line)) {
        ++number;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream fields(line);
        long long v[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 5; ++q) {
            if (!(fields >> v[q]) || v[q] < 0 || v[q] > 0xFFFFFFFFll) {
                err = "launch table line " + std::to_string(number) + ": expected five or six unsigned integers";
                return false;
            }
        }
        std::string extra;
        if (fields >> extra) {
            char* stop = nullptr;
            const long long g = std::strtoll(extra.c_str(), &stop, 10);
            if (*stop != '\0' || g < 0 || g > 0xFFFFFFFFll) {
                err = "launch table line " + std::to_string(number) + ": expected five or six unsigned integers";
                return false;
            }
            v[5] = g;
        }
        if (fields >> extra) {
            err = "launch table line " + std::to_string(number) + ": trailing text";
            return false;
        }
        if (v[1] == 0) {
            err = "launch table line " + std::to_string(number) + ": zero frames";
            return false;
        }
        table[{(uint32_t)v[0], (uint32_t)v[1]}] =
            LaunchParams{(uint32_t)v[2], (uint32_t)v[3], (uint32_t)v[4], (uint32_t)v[5]};
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    g_table = std::move(table);
    ++g_version;
    return true;
}

// The table named by MVDR_LAUNCH_TABLE, read once before the first lookup.
void load_from_environment() {
    const char* path = std::getenv("MVDR_LAUNCH_TABLE");
    if (!path || !*path) return;
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "mvdr: MVDR_LAUNCH_TABLE %s: cannot read; using the built-in launch rules\n", path);
        return;
    }
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::string err;
    if (!store_table_text(text, err))
        std::fprintf(stderr, "mvdr: MVDR_LAUNCH_TABLE %s: %s; using the built-in launch rules\n", path, err.c_str());
}

}  // namespace

uint32_t node_bucket(uint32_t nodes) { return (nodes + 1023) / 1024; }

LaunchParams launch_params(uint32_t nodes, uint32_t frames) {
    std::call_once(g_env_once, load_from_environment);
    std::lock_guard<std::mutex> lock(g_mutex);
    const auto it = g_table.find({node_bucket(nodes), frames});
    return it == g_table.end() ? LaunchParams{0, 0, 0, 0} : it->second;
}

void set_launch_params(uint32_t nodes, uint32_t frames, const LaunchParams& p) {
    std::call_once(g_env_once, load_from_environment);
    std::lock_guard<std::mutex> lock(g_mutex);
    g_table[{node_bucket(nodes), frames}] = p;
    ++g_version;
}

void clear_launch_params() {
    std::call_once(g_env_once, load_from_environment);
    std::lock_guard<std::mutex> lock(g_mutex);
    g_table.clear();
    ++g_version;
}

uint64_t launch_table_version() {
    std::call_once(g_env_once, load_from_environment);
    std::lock_guard<std::mutex> lock(g_mutex);
    return g_version;
}

std::string launch_table_text() {
    std::call_once(g_env_once, load_from_environment);
    std::lock_guard<std::mutex> lock(g_mutex);
    std::ostringstream out;
    out << "# mvdr launch table: node_bucket(1024 nodes) frames cg_workgroups eval_workgroups transform_workgroups "
           "gram_rows\n";
    for (const auto& kv : g_table)
        out << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.cg_workgroups << ' '
            << kv.second.eval_workgroups << ' ' << kv.second.transform_workgroups << ' ' << kv.second.gram_rows << '\n';
    return out.str();
}

bool load_launch_table_text(const std::string& text, std::string& err) {
    std::call_once(g_env_once, load_from_environment);
    return store_table_text(text, err);
}

}  // namespace mvdr

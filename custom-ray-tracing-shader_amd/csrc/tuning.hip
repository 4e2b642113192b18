// tuning.hip - the launch autotuner's staged search and its JSON cache (host code).
#include "tuning.h"
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

namespace rt {

const TunedField kTunedFields[8] = {
    {"bvh_width", &LaunchOptions::bvhWidth},       {"build_max_leaf", &LaunchOptions::buildMaxLeaf},
    {"blocks_per_cu", &LaunchOptions::blocksPerCU}, {"short_stack", &LaunchOptions::shortStack},
    {"while_while", &LaunchOptions::whileWhile},   {"refill", &LaunchOptions::refill},
    {"pipelines", &LaunchOptions::pipelines},      {"tail_bounce", &LaunchOptions::tailBounce}};

namespace {

std::mutex gCacheMutex;
using Entry = std::map<std::string, double>;
using Cache = std::map<std::string, Entry>;

int log2Class(long long n) { return (int)std::lround(std::log2((double)(n > 1 ? n : 1))); }

// Reader for the cache's two-level JSON object {"key": {"name": number, ...}, ...}; any other
// shape reads as an empty cache.
struct Reader {
    const std::string& t;
    size_t i = 0;
    explicit Reader(const std::string& text) : t(text) {}
    void ws() { while (i < t.size() && std::isspace((unsigned char)t[i])) i++; }
    bool eat(char c) { ws(); if (i < t.size() && t[i] == c) { i++; return true; } return false; }
    bool str(std::string& out) {
        ws();
        if (i >= t.size() || t[i] != '"') return false;
        out.clear();
        for (i++; i < t.size() && t[i] != '"'; i++) {
            if (t[i] == '\\' && i + 1 < t.size()) i++;
            out.push_back(t[i]);
        }
        if (i >= t.size()) return false;
        i++;
        return true;
    }
    bool num(double& out) {
        ws();
        const char* b = t.c_str() + i;
        char* e = nullptr;
        out = std::strtod(b, &e);
        if (e == b) return false;
        i += (size_t)(e - b);
        return true;
    }
    bool entry(Entry& e) {
        if (!eat('{')) return false;
        if (eat('}')) return true;
        do {
            std::string k;
            double v;
            if (!str(k) || !eat(':') || !num(v)) return false;
            e[k] = v;
        } while (eat(','));
        return eat('}');
    }
    bool cache(Cache& c) {
        if (!eat('{')) return false;
        if (eat('}')) return true;
        do {
            std::string k;
            Entry e;
            if (!str(k) || !eat(':') || !entry(e)) return false;
            c[k] = e;
        } while (eat(','));
        return eat('}');
    }
};

Cache readCache(const std::string& path) {
    Cache c;
    std::ifstream in(path);
    if (!in) return c;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    Reader r(text);
    if (!r.cache(c)) c.clear();
    return c;
}

bool writeCache(const std::string& path, const Cache& c) {
    for (size_t k = path.find('/', 1); k != std::string::npos; k = path.find('/', k + 1))
        (void)mkdir(path.substr(0, k).c_str(), 0755);   // each directory of the path, as mkdir -p
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream out(tmp);
        if (!out) return false;
        out << "{";
        bool first = true;
        for (const auto& kv : c) {
            out << (first ? "\n" : ",\n") << " \"" << kv.first << "\": {";
            first = false;
            bool f2 = true;
            for (const auto& fv : kv.second) {
                out << (f2 ? "" : ", ") << "\"" << fv.first << "\": ";
                if (fv.second == std::floor(fv.second) && std::fabs(fv.second) < 1e15) out << (long long)fv.second;
                else out << fv.second;
                f2 = false;
            }
            out << "}";
        }
        out << "\n}\n";
        if (!out) return false;
    }
    return std::rename(tmp.c_str(), path.c_str()) == 0;
}

}  // namespace

std::string tuneCachePath() {
    if (const char* env = std::getenv("RT_AUTOTUNE_CACHE")) return env;
    const char* home = std::getenv("HOME");
    return std::string(home ? home : ".") + "/.cache/rt_amd/autotune.json";
}

std::string sessionTunePrefix(long long triangles, int width, int height) {
    return "tris2^" + std::to_string(log2Class(triangles)) + "-px" + std::to_string(width) + "x" + std::to_string(height) +
           "-b";
}
std::string sessionTuneKey(long long triangles, int width, int height, int maxBounces) {
    return sessionTunePrefix(triangles, width, height) + std::to_string(maxBounces);
}
std::string batchTuneKey(int views, long long triangles, long long pixels, int maxBounces) {
    return "batch-v" + std::to_string(views) + "-tris2^" + std::to_string(log2Class(triangles)) + "-px2^" +
           std::to_string(log2Class(pixels)) + "-b" + std::to_string(maxBounces);
}

bool tuneLookup(const std::string& key, bool prefix, LaunchOptions& out) {
    std::lock_guard<std::mutex> lock(gCacheMutex);
    const Cache c = readCache(tuneCachePath());
    for (const auto& kv : c) {
        if (prefix ? kv.first.compare(0, key.size(), key) != 0 : kv.first != key) continue;
        LaunchOptions o;
        for (const TunedField& f : kTunedFields) {
            auto it = kv.second.find(f.name);
            if (it != kv.second.end()) o.*(f.field) = (int)it->second;
        }
        auto it = kv.second.find("mrays_per_s");
        if (it != kv.second.end()) o.mraysPerSecond = it->second;
        out = o;
        return true;
    }
    return false;
}

bool tuneStore(const std::string& key, const LaunchOptions& o) {
    std::lock_guard<std::mutex> lock(gCacheMutex);
    const std::string path = tuneCachePath();
    Cache c = readCache(path);
    Entry e;
    for (const TunedField& f : kTunedFields)
        if (o.*(f.field) >= 0) e[f.name] = o.*(f.field);
    e["mrays_per_s"] = std::round(o.mraysPerSecond * 100) / 100;
    c[key] = e;
    return writeCache(path, c);
}

LaunchOptions tuneSearch(const TuneSpace& sp, int maxBounces, const std::function<double(const LaunchOptions&)>& measure,
                         int* trials) {
    LaunchOptions best;
    best.mraysPerSecond = -1;
    int n = 0;
    auto trial = [&](LaunchOptions c) {
        c.mraysPerSecond = measure(c);
        n++;
        if (c.mraysPerSecond > best.mraysPerSecond) best = c;
    };
    const bool tails = !sp.tails.empty();
    // Stage 1: every layout x workgroups x stack x loop, at two pipelines and no path completion.
    for (int w : sp.widths)
        for (int leaf : sp.leaves)
            for (int b : sp.blocks)
                for (int st : sp.stacks)
                    for (const auto& lp : sp.loops) {
                        LaunchOptions c;
                        c.bvhWidth = w; c.buildMaxLeaf = leaf; c.blocksPerCU = b; c.shortStack = st;
                        c.whileWhile = lp.first; c.refill = lp.second; c.pipelines = 2; c.tailBounce = tails ? 0 : -1;
                        trial(c);
                    }
    if (best.mraysPerSecond <= 0) { if (trials) *trials = n; return best; }
    // Stage 2..: one knob at a time around the best; each depends on what shares the GPU.
    LaunchOptions base = best;
    for (int p : sp.pipelines)
        if (p != base.pipelines) { LaunchOptions c = base; c.pipelines = p; trial(c); }
    base = best;
    for (int b : sp.blocks)
        if (b != base.blocksPerCU) { LaunchOptions c = base; c.blocksPerCU = b; trial(c); }
    base = best;
    for (int t : sp.tails)
        if (t != base.tailBounce && t < maxBounces) { LaunchOptions c = base; c.tailBounce = t; trial(c); }
    base = best;
    for (int st : sp.stacks)
        for (const auto& lp : sp.loops)
            if (st != base.shortStack || lp.first != base.whileWhile || lp.second != base.refill) {
                LaunchOptions c = base;
                c.shortStack = st; c.whileWhile = lp.first; c.refill = lp.second;
                trial(c);
            }
    if (trials) *trials = n;
    return best;
}

}  // namespace rt

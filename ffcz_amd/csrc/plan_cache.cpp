// FFT plan cache of the autotuner (plan_cache.hpp).
#include "plan_cache.hpp"

#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <vector>

namespace ffcz {
namespace {

// Identity of this compile of the library (its compile date and time, one
// whitespace-free token): tile variants are compile-time kernel shapes, so a
// plan only holds for the build that timed it.
const std::string& build_id() {
    static const std::string id = [] {
        std::string s = __DATE__ "_" __TIME__;
        for (char& c : s)
            if (c == ' ' || c == '\t') c = '_';
        return s;
    }();
    return id;
}

struct Entry {
    std::string device;
    PlanKey k;
    TileVars v;
};

bool same(const Entry& e, const std::string& dev, const PlanKey& k) {
    return e.device == dev && e.k.bytes == k.bytes && e.k.nx == k.nx && e.k.ny == k.ny && e.k.nz == k.nz &&
           e.k.prec == k.prec && e.k.mode == k.mode;
}

std::mutex g_mu;
std::vector<Entry> g_process;  // tuned or adopted in this process
std::vector<Entry> g_disk;     // the file's plans for this compile, read once
bool g_disk_read = false;
bool g_path_set = false;
std::string g_path;

std::string default_path() {
    const char* env = std::getenv("FFCZ_PLAN_CACHE");
    if (env) return (std::strcmp(env, "off") == 0) ? std::string() : std::string(env);
    const char* xdg = std::getenv("XDG_CACHE_HOME");
    if (xdg && *xdg) return std::string(xdg) + "/ffcz/plans.txt";
    const char* home = std::getenv("HOME");
    if (home && *home) return std::string(home) + "/.cache/ffcz/plans.txt";
    return std::string();
}

const std::string& path_locked() {
    if (!g_path_set) {
        g_path = default_path();
        g_path_set = true;
    }
    return g_path;
}

// "<name>|<arch>|<compute units>" of the current device, spaces as '_'
std::string device_key() {
    int dev = 0;
    hipDeviceProp_t p{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    std::string s = std::string(p.name) + "|" + p.gcnArchName + "|" + std::to_string(p.multiProcessorCount);
    for (char& c : s)
        if (c == ' ' || c == '\t' || c == '\n') c = '_';
    return s;
}

std::string format(const Entry& e) {
    std::ostringstream o;
    o << build_id() << ' ' << e.device << ' ' << e.k.bytes << ' ' << e.k.nx << ' ' << e.k.ny << ' ' << e.k.nz << ' '
      << e.k.prec << ' ' << e.k.mode << ' ' << e.v.zf << ' ' << e.v.zi << ' ' << e.v.y << ' ' << e.v.x << ' '
      << e.v.czf << ' ' << e.v.cy << ' ' << e.v.cx << ' ' << e.v.cpair << '\n';
    return o.str();
}

// One line of the file: true for a well-formed plan of this compile.
bool parse(const std::string& line, Entry& e) {
    std::istringstream in(line);
    std::string build;
    if (!(in >> build >> e.device >> e.k.bytes >> e.k.nx >> e.k.ny >> e.k.nz >> e.k.prec >> e.k.mode >> e.v.zf >>
          e.v.zi >> e.v.y >> e.v.x >> e.v.czf >> e.v.cy >> e.v.cx >> e.v.cpair))
        return false;
    if (build != build_id()) return false;
    const int* p = &e.v.zf;
    for (int i = 0; i < 7; ++i)
        if (p[i] != 0 && p[i] != 1 && !((i == 2 || i == 3) && p[i] == 2)) return false;
    return e.v.cpair == 0 || e.v.cpair == 1;
}

std::string read_all(int fd) {
    std::string s;
    char buf[4096];
    if (lseek(fd, 0, SEEK_SET) < 0) return s;
    for (;;) {
        const ssize_t n = read(fd, buf, sizeof(buf));
        if (n <= 0) break;
        s.append(buf, (size_t)n);
    }
    return s;
}

void read_disk_locked() {
    if (g_disk_read) return;
    g_disk_read = true;
    const std::string& path = path_locked();
    if (path.empty()) return;
    const int fd = open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) return;
    if (flock(fd, LOCK_SH) == 0) {
        std::istringstream in(read_all(fd));
        std::string line;
        Entry e;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#' || !parse(line, e)) continue;
            bool replaced = false;
            for (auto& d : g_disk)
                if (same(d, e.device, e.k)) {
                    d.v = e.v;
                    replaced = true;
                }
            if (!replaced) g_disk.push_back(e);
        }
        flock(fd, LOCK_UN);
    }
    close(fd);
}

void mkdirs(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') {
            const std::string d = path.substr(0, i);
            (void)mkdir(d.c_str(), 0755);
        }
}

// Merges e into the file under an exclusive lock: the lines of other keys
// (and of other builds) are kept, this key's line is replaced.
void write_disk_locked(const Entry& e) {
    const std::string& path = path_locked();
    if (path.empty()) return;
    mkdirs(path);
    const int fd = open(path.c_str(), O_RDWR | O_CREAT | O_CLOEXEC, 0644);
    if (fd < 0) return;
    if (flock(fd, LOCK_EX) == 0) {
        std::istringstream in(read_all(fd));
        std::string out = "# ffcz FFT plan cache: build device bytes nx ny nz fft mode | zf zi y x czf cy cx cpair\n";
        std::string line;
        Entry o;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            if (parse(line, o) && same(o, e.device, e.k)) continue;
            out += line + '\n';
        }
        out += format(e);
        if (lseek(fd, 0, SEEK_SET) == 0 && ftruncate(fd, 0) == 0) {
            size_t off = 0;
            while (off < out.size()) {
                const ssize_t n = write(fd, out.data() + off, out.size() - off);
                if (n <= 0) break;
                off += (size_t)n;
            }
        }
        flock(fd, LOCK_UN);
    }
    close(fd);
}

}  // namespace

int plan_lookup(const PlanKey& k, TileVars& v) {
    const std::string dev = device_key();
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& e : g_process)
        if (same(e, dev, k)) {
            v = e.v;
            return TUNE_PROCESS;
        }
    read_disk_locked();
    for (const auto& e : g_disk)
        if (same(e, dev, k)) {
            v = e.v;
            g_process.push_back(e);
            return TUNE_DISK;
        }
    return TUNE_NONE;
}

void plan_store(const PlanKey& k, const TileVars& v) {
    Entry e{device_key(), k, v};
    std::lock_guard<std::mutex> lk(g_mu);
    bool found = false;
    for (auto& p : g_process)
        if (same(p, e.device, k)) {
            p.v = v;
            found = true;
        }
    if (!found) g_process.push_back(e);
    write_disk_locked(e);
}

std::string plan_cache_path() {
    std::lock_guard<std::mutex> lk(g_mu);
    return path_locked();
}

void set_plan_cache_path(const char* path) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_path = path ? std::string(path) : std::string();
    g_path_set = true;
    g_disk.clear();
    g_disk_read = false;
}

}  // namespace ffcz

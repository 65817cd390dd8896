// Host placement near the GPU (see pm_placement.h).
#include "pm_placement.h"

#include <sched.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace {

int read_int(const std::string& path) {
    std::ifstream f(path);
    int v = -1;
    if (!(f >> v)) return -1;
    return v;
}

std::string read_line(const std::string& path) {
    std::ifstream f(path);
    std::string s;
    std::getline(f, s);
    return s;
}

bool same_core(const PmCpu& a, const PmCpu& b) {
    if (a.cpu == b.cpu) return true;
    return a.package >= 0 && a.core >= 0 && a.package == b.package && a.core == b.core;
}

}  // namespace

std::vector<int> pm_parse_cpulist(const std::string& list) {
    std::vector<int> cpus;
    std::stringstream parts(list);
    std::string part;
    while (std::getline(parts, part, ',')) {
        int first = -1, last = -1;
        const int got = std::sscanf(part.c_str(), "%d-%d", &first, &last);
        if (got < 1 || first < 0) continue;
        if (got == 1) last = first;
        if (last < first || last >= CPU_SETSIZE) continue;
        for (int c = first; c <= last; ++c) cpus.push_back(c);
    }
    return cpus;
}

int pm_pci_numa_node(const std::string& bus_id) {
    unsigned domain = 0, bus = 0, dev = 0, fn = 0;
    if (std::sscanf(bus_id.c_str(), "%x:%x:%x.%x", &domain, &bus, &dev, &fn) != 4) return -1;
    char canonical[32];
    std::snprintf(canonical, sizeof(canonical), "%04x:%02x:%02x.%x", domain, bus, dev, fn);
    return read_int(std::string("/sys/bus/pci/devices/") + canonical + "/numa_node");
}

std::vector<PmCpu> pm_allowed_cpus(int node) {
    cpu_set_t allowed;
    CPU_ZERO(&allowed);
    std::vector<PmCpu> out;
    if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return out;
    std::vector<int> candidates;
    if (node >= 0) candidates = pm_parse_cpulist(read_line("/sys/devices/system/node/node" + std::to_string(node) + "/cpulist"));
    bool any = false;
    for (int c : candidates) any = any || CPU_ISSET(c, &allowed);
    if (!any) {
        candidates.clear();
        for (int c = 0; c < CPU_SETSIZE; ++c) candidates.push_back(c);
    }
    for (int c : candidates) {
        if (!CPU_ISSET(c, &allowed)) continue;
        const std::string topo = "/sys/devices/system/cpu/cpu" + std::to_string(c) + "/topology/";
        out.push_back({c, read_int(topo + "physical_package_id"), read_int(topo + "core_id")});
    }
    return out;
}

std::vector<int> pm_worker_order(const std::vector<PmCpu>& cpus, int caller) {
    PmCpu me{caller, -1, -1};
    for (const PmCpu& c : cpus)
        if (c.cpu == caller) me = c;
    std::vector<PmCpu> picked;
    const auto taken = [&](const PmCpu& c) {
        for (const PmCpu& p : picked)
            if (p.cpu == c.cpu) return true;
        return false;
    };
    // one CPU per physical core, the caller's core left out
    for (const PmCpu& c : cpus) {
        if (same_core(c, me)) continue;
        bool core_used = false;
        for (const PmCpu& p : picked) core_used = core_used || same_core(p, c);
        if (!core_used) picked.push_back(c);
    }
    // second hyperthreads of those cores
    for (const PmCpu& c : cpus)
        if (!same_core(c, me) && !taken(c)) picked.push_back(c);
    // the caller's core's other hyperthreads, then the caller's own CPU
    for (const PmCpu& c : cpus)
        if (c.cpu != caller && !taken(c)) picked.push_back(c);
    for (const PmCpu& c : cpus)
        if (!taken(c)) picked.push_back(c);
    std::vector<int> order;
    for (const PmCpu& c : picked) order.push_back(c.cpu);
    return order;
}

bool pm_pin_thread(const std::vector<int>& cpus) {
    cpu_set_t set;
    CPU_ZERO(&set);
    bool any = false;
    for (int c : cpus) {
        if (c < 0 || c >= CPU_SETSIZE) continue;
        CPU_SET(c, &set);
        any = true;
    }
    return any && sched_setaffinity(0, sizeof(set), &set) == 0;
}

bool pm_placement_enabled() {
    const char* v = std::getenv("PM_HOST_PLACEMENT");
    return !(v && v[0] == '0' && v[1] == '\0');
}

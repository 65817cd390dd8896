// CPU check of the host placement helpers (csrc/pm_placement.cpp): cpulist
// parsing, the worker order over synthetic topologies, and pinning the calling
// thread on this machine.  Prints "ok" and exits 0 when every check passes.
#include <sched.h>

#include <cstdio>
#include <vector>

#include "pm_placement.h"

static int failures = 0;

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

static bool same(const std::vector<int>& a, const std::vector<int>& b) { return a == b; }

int main() {
    expect(same(pm_parse_cpulist("0-3,8,10-11"), {0, 1, 2, 3, 8, 10, 11}), "cpulist ranges");
    expect(same(pm_parse_cpulist("5"), {5}), "cpulist single");
    expect(pm_parse_cpulist("").empty(), "cpulist empty");
    expect(same(pm_parse_cpulist("x,2-1,4"), {4}), "cpulist malformed parts skipped");

    // two packages x two cores x two hyperthreads: cpu = thread*4 + package*2 + core
    std::vector<PmCpu> topo;
    for (int cpu = 0; cpu < 8; ++cpu) topo.push_back({cpu, (cpu >> 1) & 1, cpu & 1});
    // caller on cpu 0 (package 0, core 0; its sibling is cpu 4)
    expect(same(pm_worker_order(topo, 0), {1, 2, 3, 5, 6, 7, 4, 0}), "order: cores first, caller's core last");
    // caller on a second hyperthread
    expect(same(pm_worker_order(topo, 7), {0, 1, 2, 4, 5, 6, 3, 7}), "order: caller on a sibling");
    // unknown topology: every CPU its own core, caller last
    std::vector<PmCpu> flat = {{2, -1, -1}, {3, -1, -1}, {9, -1, -1}};
    expect(same(pm_worker_order(flat, 3), {2, 9, 3}), "order: unknown topology");
    // caller outside the set
    expect(same(pm_worker_order(flat, 40), {2, 3, 9}), "order: caller not in the set");
    expect(pm_worker_order({}, 0).empty(), "order: empty set");

    expect(pm_pci_numa_node("not a bus id") == -1, "bus id parse failure");
    expect(pm_pci_numa_node("ffff:ff:1f.7") == -1, "absent device");

    // live: the allowed CPUs of this process, pin to the first, then restore
    cpu_set_t before;
    sched_getaffinity(0, sizeof(before), &before);
    const std::vector<PmCpu> mine = pm_allowed_cpus(-1);
    expect(!mine.empty(), "allowed cpus listed");
    expect(int(mine.size()) == CPU_COUNT(&before), "allowed cpus match the affinity mask");
    if (!mine.empty()) {
        expect(pm_pin_thread({mine[0].cpu}), "pin to one cpu");
        cpu_set_t now;
        sched_getaffinity(0, sizeof(now), &now);
        expect(CPU_COUNT(&now) == 1 && CPU_ISSET(mine[0].cpu, &now), "pinned mask");
        expect(sched_getcpu() == mine[0].cpu, "running on the pinned cpu");
    }
    expect(!pm_pin_thread({}), "empty set refused");
    sched_setaffinity(0, sizeof(before), &before);
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}

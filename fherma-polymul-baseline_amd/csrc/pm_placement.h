// Host placement near the GPU (role of the reference's host_affinity.h:14-46
// and host_cpu_topology.h:15-36): the CPUs of the device's NUMA node that this
// process may run on, and an order for worker threads that takes one logical
// CPU per physical core before any second hyperthread and keeps off the
// calling thread's core.  Linux sysfs; missing topology leaves placement as it
// is.  Host-only code (no HIP calls), so it builds and tests without a GPU.
#pragma once

#include <string>
#include <vector>

struct PmCpu {
    int cpu = -1;
    int package = -1;  // physical package id, -1 if unknown
    int core = -1;     // core id within the package, -1 if unknown
};

// "0-3,8,10-11" -> {0,1,2,3,8,10,11}; malformed parts are skipped.
std::vector<int> pm_parse_cpulist(const std::string& list);

// NUMA node of the PCI device with bus id "dddd:bb:dd.f" (hex, any case and
// width, as hipDeviceGetPCIBusId writes it); -1 when unknown.
int pm_pci_numa_node(const std::string& bus_id);

// The CPUs this thread may run on (sched_getaffinity) that belong to NUMA
// node `node` (all of them when node < 0 or the node lists none), with their
// package and core ids.
std::vector<PmCpu> pm_allowed_cpus(int node);

// Order for worker threads over `cpus`: one CPU of each physical core other
// than the caller's, then the remaining hyperthreads of those cores, then the
// caller's core's other hyperthreads; `caller` itself comes last.  Unknown
// topology counts every CPU as its own core.
std::vector<int> pm_worker_order(const std::vector<PmCpu>& cpus, int caller);

// Restricts the calling thread to `cpus`; false (and no change) when the set
// is empty or the kernel refuses it.
bool pm_pin_thread(const std::vector<int>& cpus);

// Whether host placement is enabled: PM_HOST_PLACEMENT=0 in the environment
// turns it off (worker pinning and the drop-in's caller placement).
bool pm_placement_enabled();

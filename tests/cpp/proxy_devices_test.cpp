// Self-checking test of the device proxies of the C++ proxy ABI
// (include/cudaGSEA_wrapper.cuh) against the behaviour of
// cudaGSEA/src/cudaGSEA_wrapper.cu:87-138:
//   listCudaDevices_proxy()  one string per device,
//                            "device <id>: <name> with <GiB> GiB RAM",
//                            GiB = total device memory rounded up;
//   getCudaDevice_proxy()    the active device;
//   setCudaDevice_proxy(id)  selects id and returns it; an id that cannot be
//                            selected keeps the active device, returns it and
//                            prints "WARNING: invalid device id, fallback to
//                            device <active>." on stdout.
// The expected strings are built from hipGetDeviceProperties directly.
// Exit status 0 when every check holds; prints one line per failure.
#include <hip/hip_runtime_api.h>

#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cudaGSEA_wrapper.cuh"

static int failures = 0;

static void expect(bool ok, const std::string &what) {
    if (!ok) {
        ++failures;
        std::cerr << "FAIL: " << what << std::endl;
    }
}

// runs f with std::cout captured; returns what it printed
template <class F>
static std::string captured(F f) {
    std::ostringstream buf;
    std::streambuf *old = std::cout.rdbuf(buf.rdbuf());
    f();
    std::cout.rdbuf(old);
    return buf.str();
}

int main() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;

    const std::vector<std::string> list = listCudaDevices_proxy();
    expect((int)list.size() == n, "listCudaDevices_proxy returns one entry per device (" +
                                      std::to_string(list.size()) + " vs " +
                                      std::to_string(n) + ")");
    for (int i = 0; i < n && i < (int)list.size(); ++i) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, i) != hipSuccess) {
            expect(false, "hipGetDeviceProperties " + std::to_string(i));
            continue;
        }
        const size_t gib = (prop.totalGlobalMem + ((size_t)1 << 30) - 1) >> 30;
        const std::string name = prop.name[0] ? prop.name : prop.gcnArchName;
        const std::string want = "device " + std::to_string(i) + ": " + name + " with " +
                                 std::to_string(gib) + " GiB RAM";
        expect(list[i] == want, "device string '" + list[i] + "' != '" + want + "'");
        std::cout << list[i] << std::endl;
    }

    // every valid id is selected silently
    for (int i = 0; i < n; ++i) {
        int got = -1;
        const std::string out = captured([&] { got = setCudaDevice_proxy(i); });
        expect(got == i, "setCudaDevice_proxy(" + std::to_string(i) + ") returns it");
        expect(getCudaDevice_proxy() == i, "getCudaDevice_proxy after set");
        expect(out.empty(), "no output for a valid id, got '" + out + "'");
    }

    // ids that cannot be selected fall back to the active device
    int active = 0;
    if (n > 0 && hipGetDevice(&active) != hipSuccess) active = 0;
    for (int bad : {n, n + 7, -1}) {
        int got = -2;
        const std::string out = captured([&] { got = setCudaDevice_proxy(bad); });
        // (the warning compares the id with the device left active; with no
        // device at all, id 0 and the default 0 agree)
        const std::string want =
            active == bad ? std::string()
                          : "WARNING: invalid device id, fallback to device " +
                                std::to_string(active) + ".\n";
        expect(got == active, "setCudaDevice_proxy(" + std::to_string(bad) +
                                  ") returns the active device");
        expect(out == want, "fallback warning '" + out + "' != '" + want + "'");
    }
    if (n > 0) expect(getCudaDevice_proxy() == active, "active device unchanged");

    std::cout << (failures ? "FAILED " : "ok ") << n << " device(s), " << failures
              << " failure(s)" << std::endl;
    return failures ? 1 : 0;
}

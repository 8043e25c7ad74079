// spdk: command-line driver, "spdk -i file.athinput [block/name=value ...]"
// (src/main.cpp:86-257). Outputs go to $SPD_OUTPUT_DIR (default ./output).
#include <cstdlib>
#include "../../../include/spdk.h"
int main(int argc, char** argv) {
    const char* od = std::getenv("SPD_OUTPUT_DIR");
    spdk_run_result r;
    int rc = spdk_run_argv(argc - 1, (const char* const*)(argv + 1), od ? od : "output", &r);
    return rc == 0 ? 0 : 1;
}

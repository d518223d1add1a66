// soapBubble: the reference's command-line driver (kernel/main.cu:4-56) over
// libsoapfilm.so.
//   soapBubble config.txt [--precision fp64|mixed] [--advection bimocq|plain]
//                         [--air airflow.exr] [--viscous on|off]
#include <cstring>
#include <iostream>
#include <string>

#include "bubble.h"

static int usage() {
    std::cout << "Please provide the path to config.txt as an argument." << std::endl;
    std::cout << "Usage example: ./soapBubble ../config.txt [--precision fp64|mixed] "
                 "[--advection bimocq|plain] [--air airflow.exr] [--viscous on|off]"
              << std::endl;
    std::cout << "Configuration file was missing, exiting." << std::endl;
    return -1;
}

int main(int argc, char** argv) {
    if (argc < 2 || (argc % 2) != 0) return usage();
    sfb::Config cfg;
    std::string err;
    if (!sfb::read_config(argv[1], cfg, err)) {
        std::cerr << err << std::endl;
        return 2;
    }
    try {
        sfb::Bubble bubble(cfg);
        for (int k = 2; k + 1 < argc; k += 2) {
            const std::string key = argv[k], val = argv[k + 1];
            if (key == "--precision" && (val == "fp64" || val == "mixed")) {
                bubble.precision = val == "mixed" ? SF_PREC_MIXED : SF_PREC_FP64;
            } else if (key == "--advection" && (val == "bimocq" || val == "plain")) {
                bubble.bimocq = val == "bimocq";
            } else if (key == "--air") {
                bubble.airImage = val;
            } else if (key == "--viscous" && (val == "on" || val == "off")) {
                bubble.viscous = val == "on";
            } else {
                std::cerr << "unknown option " << key << " " << val
                          << " (--precision fp64|mixed, --advection bimocq|plain, --air path, "
                             "--viscous on|off)"
                          << std::endl;
                return 2;
            }
        }
        const int done = bubble.run();
        std::cout << "Frames completed: " << done << std::endl;
        std::cout << "Frame loop seconds: " << bubble.loopSeconds() << std::endl;
    } catch (const sfb::Error& e) {
        std::cerr << "soapBubble: " << e.what() << std::endl;
        return 3;
    }
    return 0;
}

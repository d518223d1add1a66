// soapBubble: the reference's command-line driver (kernel/main.cu:4-56) over
// libsoapfilm.so. Usage: soapBubble config.txt [fp64|mixed] [air.exr]
#include <cstring>
#include <iostream>

#include "bubble.h"

int main(int argc, char** argv) {
    if (argc < 2 || argc > 4) {
        std::cout << "Please provide the path to config.txt as an argument." << std::endl;
        std::cout << "Usage example: ./soapBubble ../config.txt [fp64|mixed] [airflow.exr]"
                  << std::endl;
        std::cout << "Configuration file was missing, exiting." << std::endl;
        return -1;
    }
    sfb::Config cfg;
    std::string err;
    if (!sfb::read_config(argv[1], cfg, err)) {
        std::cerr << err << std::endl;
        return 2;
    }
    try {
        sfb::Bubble bubble(cfg);
        if (argc >= 3) {
            if (std::strcmp(argv[2], "mixed") == 0) {
                bubble.precision = SF_PREC_MIXED;
            } else if (std::strcmp(argv[2], "fp64") != 0) {
                std::cerr << "precision must be fp64 or mixed" << std::endl;
                return 2;
            }
        }
        if (argc == 4) bubble.airImage = argv[3];
        const int done = bubble.run();
        std::cout << "Frames completed: " << done << std::endl;
    } catch (const sfb::Error& e) {
        std::cerr << "soapBubble: " << e.what() << std::endl;
        return 3;
    }
    return 0;
}

// emfa launcher: reads the CONFIG.yml path from stdin (or argv[1]) and runs
// the solver it names, mapping exceptions to an exit-state string as
// EMLauncher/launcher.cpp:8-49 does.
#include <iostream>
#include <string>

#include "core.h"
#include "em.h"

int main(int argc, char **argv) {
    std::string config, state("OK");
    std::cout << "Input Path of Config File:";
    if (argc > 1) {
        config = argv[1];
    } else {
        std::getline(std::cin, config);
    }
    if (config.empty()) config = "CONFIG.yml";
    int rc = 0;
    try {
        solver::loadSolver(config);
    } catch (const component::ConfigInvalid &e) {
        state = std::string("[ConfigParameter Error]") + e.what();
        rc = 2;
    } catch (const component::FileError &e) {
        state = std::string("[File Error]") + e.what();
        rc = 3;
    } catch (const std::invalid_argument &e) {
        state = std::string("[InvalidArgument Error]") + e.what();
        rc = 4;
    } catch (const std::runtime_error &e) {
        state = std::string("[Runtime Error]") + e.what();
        rc = 5;
    } catch (const std::exception &e) {
        state = std::string("[Unrecognized Error]") + e.what();
        rc = 6;
    }
    std::cout << "\nExit State: " << state << "\n" << std::endl;
    return rc;
}

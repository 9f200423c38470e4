// Small C entry points of the host core, for bindings and for the tests:
// configuration lookup with the loader's semantics, the NAS reader, and a
// whole CONFIG.yml run in-process.
#include <cstring>
#include <iostream>
#include <string>

#include "core.h"
#include "em.h"

namespace {
thread_local std::string g_core_err;
int fail(int code, const std::string &msg) {
    g_core_err = msg;
    return code;
}
}  // namespace

extern "C" {

const char *emfa_core_last_error(void) { return g_core_err.c_str(); }

// Value of `key` (with its colon) after loading `src` (a path, or the text
// itself when is_text != 0).  Returns 0, 1 (missing key) or 2 (bad file).
int emfa_core_config_get(const char *src, int is_text, const char *key, char *out, int cap) {
    try {
        component::ConfigLoader L;
        std::streambuf *old = std::cout.rdbuf(nullptr);
        if (is_text)
            L.loadFromString(src);
        else
            L.load(src);
        std::cout.rdbuf(old);
        if (!L.has(key)) return fail(1, std::string("missing key ") + key);
        const std::string &v = L.raw(key);
        if ((int)v.size() + 1 > cap) return fail(2, "buffer too small");
        std::memcpy(out, v.c_str(), v.size() + 1);
        return 0;
    } catch (const std::exception &e) {
        return fail(2, e.what());
    }
}

// NAS reader (EMComponent/Mesh.cpp:261-309 dialect).  With null buffers only
// the sizes are returned.  Returns 0 or 3 (file error).
int emfa_core_read_mesh(const char *path, double *xyz, int nv_cap, int *tri, int nt_cap, int *nv, int *nt) {
    try {
        component::Mesh m;
        m.loadMeshFile(path);
        *nv = (int)m.getNodeNum();
        *nt = (int)m.getTriangleNum();
        if (!xyz || !tri) return 0;
        if (nv_cap < *nv || nt_cap < *nt) return fail(1, "buffers too small");
        std::memcpy(xyz, m.xyz().data(), sizeof(double) * m.xyz().size());
        std::memcpy(tri, m.tri().data(), sizeof(int) * m.tri().size());
        return 0;
    } catch (const std::exception &e) {
        return fail(3, e.what());
    }
}

// Runs solver::loadSolver on a config file.  Returns 0, or the launcher's
// exit codes (2 config, 3 file, 4 invalid argument, 5 runtime, 6 other).
int emfa_core_run(const char *config_path) {
    try {
        solver::loadSolver(config_path);
        return 0;
    } catch (const component::ConfigInvalid &e) {
        return fail(2, e.what());
    } catch (const component::FileError &e) {
        return fail(3, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(4, e.what());
    } catch (const std::runtime_error &e) {
        return fail(5, e.what());
    } catch (const std::exception &e) {
        return fail(6, e.what());
    }
}

}  // extern "C"

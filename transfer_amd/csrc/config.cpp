// config.cpp -- the reference's cfg grammar (Config.cpp:8-131): "key = value"
// lines, '#' comments, "[region N]" sections holding kappa / Ip.  Parse errors
// that the reference reports with `throw 0` return TR_ERR_ARG here with the
// same message.  Keys the reference leaves uninitialised when absent
// (precondition, soltype, angorder; Config.cpp:16-21) default to 0, 0, 1.
#include <cstdio>
#include <cstring>

#include "internal.h"

namespace tr {

int load_config(const char *path, Config &c) {
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(TR_ERR_ARG, std::string("failed to open file `") + path + "'");
    char buf[1024];
    int line = 0, rid = -1;
    c = Config();
    auto die = [&](const std::string &msg) {
        std::fclose(f);
        return fail(TR_ERR_ARG, std::string(path) + ":" + std::to_string(line) + ": " + msg);
    };
    while (std::fgets(buf, sizeof buf, f)) {
        line++;
        if (buf[0] == '#' || buf[0] == '\n') continue;
        if (buf[0] == '[') {
            if (std::sscanf(buf, "[region %d]", &rid) != 1)
                return die("scanf failed to parse region section");
            if (c.nreg < 0) return die("section appeared before numregs was set");
            if (rid > c.nreg || rid < 0)
                return die("region id (" + std::to_string(rid) + ") is out of bounds");
            continue;
        }
        auto strip = [&](size_t off) {
            std::string s(buf + off);
            while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
            return s;
        };
        if (!std::strncmp(buf, "kappa", 5)) {
            if (rid < 0) return die("kappa parameter outside section");
            if (std::sscanf(buf, "kappa = %lf", &c.kappa[rid]) != 1)
                return die("scanf failed to parse kappa parameter");
        } else if (!std::strncmp(buf, "Ip", 2)) {
            if (rid < 0) return die("Ip parameter outside section");
            if (std::sscanf(buf, "Ip = %lf", &c.Ip[rid]) != 1)
                return die("scanf failed to parse Ip parameter");
        } else if (!std::strncmp(buf, "meshfile = ", 11)) {
            if (rid >= 0) return die("meshfile parameter in section");
            c.meshfn = strip(11);
        } else if (!std::strncmp(buf, "outfile = ", 10)) {
            if (rid >= 0) return die("outfile parameter in section");
            c.outfn = strip(10);
        } else if (!std::strncmp(buf, "numregs", 7)) {
            if (rid >= 0) return die("numregs parameter in section");
            if (std::sscanf(buf, "numregs = %d", &c.nreg) != 1 || c.nreg < 0)
                return die("scanf failed to parse numregs parameter");
            c.kappa.assign(c.nreg + 1, 0.0);
            c.Ip.assign(c.nreg + 1, 0.0);
        } else if (!std::strncmp(buf, "angorder", 8)) {
            if (rid >= 0) return die("angorder parameter in section");
            if (std::sscanf(buf, "angorder = %d", &c.maxk) != 1)
                return die("scanf failed to parse angorder parameter");
        } else if (!std::strncmp(buf, "device", 6)) {
            if (rid >= 0) return die("device parameter in section");
            if (std::sscanf(buf, "device = %d", &c.dev) != 1)
                return die("scanf failed to parse device parameter");
        } else if (!std::strncmp(buf, "soltype", 7)) {
            if (rid >= 0) return die("soltype parameter in section");
            if (std::sscanf(buf, "soltype = %d", &c.soltype) != 1)
                return die("scanf failed to parse soltype parameter");
        } else if (!std::strncmp(buf, "dump", 4)) {
            if (rid >= 0) return die("dump parameter in section");
            if (std::sscanf(buf, "dump = %d", &c.dump) != 1)
                return die("scanf failed to parse dump parameter");
        } else if (!std::strncmp(buf, "precondition", 12)) {
            if (rid >= 0) return die("precondition parameter in section");
            if (std::sscanf(buf, "precondition = %d", &c.prec) != 1)
                return die("scanf failed to parse precondition parameter");
        }
    }
    std::fclose(f);
    return TR_OK;
}

}  // namespace tr

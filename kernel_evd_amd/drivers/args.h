// args.h -- tiny --key value command-line parser shared by the drivers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

struct Args {
    std::map<std::string, std::string> kv;
    Args(int argc, char** argv)
    {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) {
                std::fprintf(stderr, "unexpected argument: %s\n", a.c_str());
                std::exit(1);
            }
            a = a.substr(2);
            if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) kv[a] = argv[++i];
            else kv[a] = "1";
        }
    }
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    std::string str(const std::string& k, const std::string& d = "") const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    double num(const std::string& k, double d) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atof(it->second.c_str());
    }
    long long integer(const std::string& k, long long d) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atoll(it->second.c_str());
    }
    std::string need(const std::string& k) const
    {
        if (!has(k)) {
            std::fprintf(stderr, "missing required --%s\n", k.c_str());
            std::exit(1);
        }
        return str(k);
    }
};

// Map a library status onto the reference's exit codes: 2 non-positive
// degree, 3 potrf, 4 non-positive q, 5 eigensolver; 1 for anything else.
inline int exit_code(int status)
{
    if (status >= 2 && status <= 5) return status;
    return status == 0 ? 0 : 1;
}

// Named log sink taken by the DADA clients (multilog.hpp:14-56).  PSRDADA's
// multilog is not available here; messages go to stderr when FBF_LOG is set.
#ifndef PSRDADA_CPP_MULTILOG_HPP
#define PSRDADA_CPP_MULTILOG_HPP

#include <cstdio>
#include <cstdlib>
#include <string>

namespace psrdada_cpp {

class MultiLog {
public:
    explicit MultiLog(std::string const& name) : _name(name) {}
    MultiLog(MultiLog const&) = delete;
    std::string const& name() const { return _name; }
    void write(int /*priority*/, const char* msg) const
    {
        if (std::getenv("FBF_LOG")) std::fprintf(stderr, "[%s] %s\n", _name.c_str(), msg);
    }

private:
    std::string _name;
};

}  // namespace psrdada_cpp

#endif

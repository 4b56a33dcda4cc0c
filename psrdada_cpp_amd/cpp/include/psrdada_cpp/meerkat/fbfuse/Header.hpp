// ASCII DADA header access (Header.hpp:10-70).
//
// A DADA header is a text block of "KEY value" lines.  get() returns the
// value of KEY (throws if the key is missing or "unset", Header.cpp:26-39);
// set() replaces the value of an existing key or appends a new line.  This
// restates what PSRDADA's ascii_header_get/ascii_header_set do for the keys
// the pipeline uses (SAMPLE_CLOCK, SAMPLE_CLOCK_START, SYNC_TIME); PSRDADA
// itself is not available here.
#ifndef PSRDADA_CPP_MEERKAT_FBFUSE_HEADER_HPP
#define PSRDADA_CPP_MEERKAT_FBFUSE_HEADER_HPP

#include <string>

#include "psrdada_cpp/raw_bytes.hpp"

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

class Header {
public:
    explicit Header(RawBytes& header);
    ~Header();
    Header(Header const&) = delete;

    template <typename T>
    T get(char const* key);

    template <typename T>
    void set(char const* key, T value);

    // Zero the whole header block (Header.cpp:19-22).
    void purge();

    // The raw value string of key (throws if absent).
    std::string get_string(char const* key);
    void set_string(char const* key, std::string const& value);

private:
    RawBytes& _header;
};

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

#endif

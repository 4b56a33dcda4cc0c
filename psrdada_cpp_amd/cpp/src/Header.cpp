#include "psrdada_cpp/meerkat/fbfuse/Header.hpp"

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace psrdada_cpp {
namespace meerkat {
namespace fbfuse {

Header::Header(RawBytes& header) : _header(header) {}

Header::~Header() {}

void Header::purge() { std::memset(_header.ptr(), 0, _header.total_bytes()); }

namespace {
// Finds the line "KEY<space>..." at a line start; returns the offset of the
// line or npos.
std::size_t find_key(std::string const& text, std::string const& key)
{
    std::size_t pos = 0;
    while (pos < text.size()) {
        const std::size_t eol = text.find('\n', pos);
        const std::size_t end = eol == std::string::npos ? text.size() : eol;
        if (end - pos > key.size() && text.compare(pos, key.size(), key) == 0 &&
            (text[pos + key.size()] == ' ' || text[pos + key.size()] == '\t'))
            return pos;
        if (eol == std::string::npos) break;
        pos = eol + 1;
    }
    return std::string::npos;
}
}  // namespace

std::string Header::get_string(char const* key)
{
    const std::string text(_header.ptr(), strnlen(_header.ptr(), _header.total_bytes()));
    const std::size_t pos = find_key(text, key);
    if (pos == std::string::npos) throw std::runtime_error(std::string("Could not find ") + key + " key in DADA header.");
    std::size_t v = pos + std::strlen(key);
    while (v < text.size() && (text[v] == ' ' || text[v] == '\t')) ++v;
    std::size_t e = v;
    while (e < text.size() && text[e] != '\n' && text[e] != ' ' && text[e] != '\t' && text[e] != '\r') ++e;
    const std::string value = text.substr(v, e - v);
    if (value == "unset") throw std::runtime_error(std::string("The header key ") + key + " was unset");
    return value;
}

void Header::set_string(char const* key, std::string const& value)
{
    std::string text(_header.ptr(), strnlen(_header.ptr(), _header.total_bytes()));
    const std::string line = std::string(key) + " " + value + "\n";
    const std::size_t pos = find_key(text, key);
    if (pos == std::string::npos) {
        text += line;
    } else {
        const std::size_t eol = text.find('\n', pos);
        text.replace(pos, (eol == std::string::npos ? text.size() : eol + 1) - pos, line);
    }
    if (text.size() + 1 > _header.total_bytes()) throw std::runtime_error("DADA header block is full");
    std::memcpy(_header.ptr(), text.c_str(), text.size() + 1);
}

template <>
long double Header::get<long double>(char const* key)
{
    return std::strtold(get_string(key).c_str(), nullptr);
}

template <>
std::size_t Header::get<std::size_t>(char const* key)
{
    return std::strtoul(get_string(key).c_str(), nullptr, 0);
}

template <>
void Header::set<long double>(char const* key, long double value)
{
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%Lf", value);   // Header.cpp:59-63
    set_string(key, buf);
}

template <>
void Header::set<std::size_t>(char const* key, std::size_t value)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%" PRIu64, (uint64_t)value);
    set_string(key, buf);
}

}  // namespace fbfuse
}  // namespace meerkat
}  // namespace psrdada_cpp

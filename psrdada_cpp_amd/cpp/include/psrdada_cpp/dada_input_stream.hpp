// Drives a handler from a DADA ring (dada_input_stream.hpp:11-26,
// detail/dada_input_stream.cpp:21-52): per observation, attach a read
// client, hand the header to handler.init(), then every data block to
// handler(block) until end of data, a handler stop request or stop().
#ifndef PSRDADA_CPP_DADA_INPUT_STREAM_HPP
#define PSRDADA_CPP_DADA_INPUT_STREAM_HPP

#include <atomic>
#include <stdexcept>

#include "psrdada_cpp/dada_read_client.hpp"

namespace psrdada_cpp {

template <class HandlerType>
class DadaInputStream {
public:
    DadaInputStream(key_t key, MultiLog& log, HandlerType& handler)
        : _key(key), _log(log), _handler(handler), _stop(false), _running(false)
    {
    }

    void start()
    {
        if (_running) throw std::runtime_error("Stream is already running");
        _running = true;
        bool handler_stop_request = false;
        while (!_stop && !handler_stop_request) {
            DadaReadClient client(_key, _log);
            auto& header_stream = client.header_stream();
            auto& header_block = header_stream.next();
            _handler.init(header_block);
            header_stream.release();
            auto& data_stream = client.data_stream();
            while (!_stop && !handler_stop_request) {
                handler_stop_request = _handler(data_stream.next());
                data_stream.release();
                if (data_stream.at_end()) break;
            }
        }
        _running = false;
    }

    void stop() { _stop = true; }

private:
    key_t _key;
    MultiLog& _log;
    HandlerType& _handler;
    std::atomic<bool> _stop;
    bool _running;
};

// Handler that drops every block (dada_null_sink.hpp:9-16).
class NullSink {
public:
    void init(RawBytes&) {}
    bool operator()(RawBytes&) { return false; }
};

}  // namespace psrdada_cpp

#endif

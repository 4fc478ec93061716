// Logger.cc -- default handlers of the host library's logger (Logger.h).
#include "Logger.h"

#include <cstdlib>
#include <iostream>

namespace {

void toStream(std::ostream& os, const char* tag, const std::string& module, const std::string& msg, Flusher fl)
{
    os << tag << module << ": " << msg << '\n';
    if (fl == Flusher::FLUSH) os.flush();
}

}  // namespace

LoggerOutput loggerOutputDefault = {
    // FATAL and ERROR end the program (the reference's Logger.cc:194, 270)
    [](std::string m, std::string s, Flusher) {
        toStream(std::cerr, "FATAL ERROR in ", m, s, Flusher::FLUSH);
        std::exit(static_cast<int>(Log::FATAL));
    },
    [](std::string m, std::string s, Flusher) {
        toStream(std::cerr, "ERROR in ", m, s, Flusher::FLUSH);
        std::exit(static_cast<int>(Log::ERROR));
    },
    [](std::string m, std::string s, Flusher f) { toStream(std::cerr, "WARNING in ", m, s, f); },
    [](std::string, std::string s, Flusher f) {
        std::cout << s << '\n';
        if (f == Flusher::FLUSH) std::cout.flush();
    },
    [](std::string, std::string s, Flusher f) {
        std::cout << s << '\n';
        if (f == Flusher::FLUSH) std::cout.flush();
    },
    [](std::string m, std::string s, Flusher) { toStream(std::cerr, "DEBUG in ", m, s, Flusher::FLUSH); },
};

LoggerOutput* loggerOutput = &loggerOutputDefault;

const LogLevel<Log::FATAL> FATAL{};
const LogLevel<Log::ERROR> ERROR{};
const LogLevel<Log::WARN> WARN{};
const LogLevel<Log::INFO> INFO{};
const LogLevel<Log::VERBOSE> VERBOSE{};
const LogLevel<Log::DEBUG> DEBUG{};

Logger logger("MercuryKernel");

namespace logger_detail {

void unhandled(const std::string& module, const std::string& msg)
{
    throw std::runtime_error(module + ": " + msg);
}

}  // namespace logger_detail

// Logger.h -- the host library's message and error channel, with the call
// shape drivers know from the reference (Kernel/Logger.h:128-175):
//
//     logger(INFO, "inserted % particles in % s", n, t);
//     logger(ERROR, "time step % is not positive", dt);
//
// Every `%` in the format takes the next argument (printed with operator<<);
// `\%` prints a literal percent sign.  Each level calls one handler of the
// LoggerOutput that `loggerOutput` points at.  By default INFO and WARN go to
// std::cout / std::cerr and return; ERROR and FATAL print to std::cerr and end
// the program with exit status static_cast<int>(Log::ERROR) or Log::FATAL, as
// the reference's do (Logger.cc:194, 270).  A driver that wants to recover
// points loggerOutput at its own handlers, e.g. ones that throw.  An ERROR or
// FATAL handler that returns is not allowed to let the failing operation go
// on: the logger then throws std::runtime_error carrying the message.
#ifndef MERCURYDPM_HOST_LOGGER_H
#define MERCURYDPM_HOST_LOGGER_H

#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

enum class Log : signed char { FATAL = -20, ERROR = -15, WARN = -10, INFO = -5, DEFAULT = 0, VERBOSE = 5, DEBUG = 10 };

enum class Flusher { FLUSH, NO_FLUSH };

// handler signature: (module name, message, flush request)
struct LoggerOutput {
    std::function<void(std::string, std::string, Flusher)> onFatal;
    std::function<void(std::string, std::string, Flusher)> onError;
    std::function<void(std::string, std::string, Flusher)> onWarn;
    std::function<void(std::string, std::string, Flusher)> onInfo;
    std::function<void(std::string, std::string, Flusher)> onVerbose;
    std::function<void(std::string, std::string, Flusher)> onDebug;
};

// the handlers in use; never null (reset it to &loggerOutputDefault)
extern LoggerOutput* loggerOutput;
extern LoggerOutput loggerOutputDefault;

// level tags, so that levels read as single words at the call site
template <Log L>
struct LogLevel {
};
extern const LogLevel<Log::FATAL> FATAL;
extern const LogLevel<Log::ERROR> ERROR;
extern const LogLevel<Log::WARN> WARN;
extern const LogLevel<Log::INFO> INFO;
extern const LogLevel<Log::VERBOSE> VERBOSE;
extern const LogLevel<Log::DEBUG> DEBUG;

namespace logger_detail {

inline void format(std::ostringstream& os, const char* f)
{
    for (; *f; ++f) {
        if (*f == '\\' && f[1] == '%') ++f;
        os << *f;
    }
}

template <typename T, typename... Rest>
void format(std::ostringstream& os, const char* f, T&& first, Rest&&... rest)
{
    for (; *f; ++f) {
        if (*f == '\\' && f[1] == '%') {
            os << '%';
            ++f;
        } else if (*f == '%') {
            os << std::forward<T>(first);
            format(os, f + 1, std::forward<Rest>(rest)...);
            return;
        } else {
            os << *f;
        }
    }
    // more arguments than placeholders: the rest follow, space-separated
    os << ' ' << std::forward<T>(first);
    format(os, "", std::forward<Rest>(rest)...);
}

[[noreturn]] void unhandled(const std::string& module, const std::string& msg);

}  // namespace logger_detail

class Logger {
public:
    explicit Logger(std::string module) : module_(std::move(module)) {}

    template <Log L, typename... Args>
    void operator()(LogLevel<L>, const char* fmt, Args&&... args) const
    {
        if (L > level_) return;
        std::ostringstream os;
        os.precision(15);
        logger_detail::format(os, fmt, std::forward<Args>(args)...);
        const std::string msg = os.str();
        switch (L) {
        case Log::FATAL:
            call(loggerOutput->onFatal, msg, Flusher::FLUSH);
            logger_detail::unhandled(module_, msg);
        case Log::ERROR:
            call(loggerOutput->onError, msg, Flusher::FLUSH);
            logger_detail::unhandled(module_, msg);
        case Log::WARN:
            call(loggerOutput->onWarn, msg, Flusher::FLUSH);
            break;
        case Log::INFO:
        case Log::DEFAULT:
            call(loggerOutput->onInfo, msg, Flusher::NO_FLUSH);
            break;
        case Log::VERBOSE:
            call(loggerOutput->onVerbose, msg, Flusher::NO_FLUSH);
            break;
        case Log::DEBUG:
            call(loggerOutput->onDebug, msg, Flusher::FLUSH);
            break;
        }
    }

    // messages above this level are dropped (VERBOSE and DEBUG by default)
    void setLevel(Log level) { level_ = level; }
    Log getLevel() const { return level_; }

private:
    void call(const std::function<void(std::string, std::string, Flusher)>& f, const std::string& msg,
              Flusher fl) const
    {
        if (f) f(module_, msg, fl);
    }

    std::string module_;
    Log level_ = Log::DEFAULT;
};

// the host library's logger (module "MercuryKernel")
extern Logger logger;

#endif

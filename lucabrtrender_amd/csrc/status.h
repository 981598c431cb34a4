// The error codes and status of every host call (the reference's
// lrt::ErrorCode and Result, modules/core/include/lrt/core/Error.h:14-53,
// Result.h:14-95): never thrown, returned.
#pragma once

#include <string>
#include <utility>

namespace lrt {

enum class ErrorCode : int {   // modules/core/include/lrt/core/Error.h:14-26
    Ok = 0,
    Unknown,
    NotFound,
    InvalidArgument,
    Unsupported,
    IoFailure,
    DeviceFailure,
    ShaderFailure,
    PluginFailure,
    OutOfMemory,
    Cancelled,
    InternalError,
};

struct Status {
    ErrorCode code = ErrorCode::Ok;
    std::string message;
    [[nodiscard]] bool ok() const noexcept { return code == ErrorCode::Ok; }
    static Status error(ErrorCode c, std::string m) { return Status{c, std::move(m)}; }
};

}   // namespace lrt

// Last-error slot shared by the device and host entry points of libmikroptim_hip.so
// (one per calling thread; mko_last_error() returns it).
#pragma once
#include <string>

namespace mko_detail {
inline thread_local std::string last_error;
}

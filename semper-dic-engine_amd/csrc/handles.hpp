// The C ABI's handle types (include/semper_c.h declares them opaque), shared
// by the drop-in entry points and the kernel-level interface of
// libsemper_dev.so, which can reach a session's engine settings.
#pragma once

#include <atomic>

#include "engine.hpp"
#include "incremental.hpp"

struct semper_engine {
    sdic::Engine eng;
    sdic::Cancel cancel;
    std::atomic<bool> busy{false};
};

struct semper_incremental {
    sdic::IncrementalSession session;
    sdic::Cancel cancel;
    int width = 0, height = 0;
    std::atomic<bool> busy{false};
};

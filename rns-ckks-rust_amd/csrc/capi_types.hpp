// Opaque C-ABI handle types shared by the per-op entry points (capi.hip) and the batched
// executor (batch.hip).
#pragma once
#include "pipeline.hpp"

struct rck_context {
    rck::Context* c;
};

struct rck_ksk : rck::KeyDev {
    rck_context* owner = nullptr;
};

// egsgpu_engine.h -- host-side engine objects behind the C ABI handles.
#pragma once
#include "egs_tables.h"

struct egsgpu_tables {
    egsgpu::Blob blob;
};

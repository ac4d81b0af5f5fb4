// Opaque handles of the C ABI shared by the prover and the executor.
#pragma once
#include <memory>

#include "config.h"
#include "transcript.h"

struct whir_config {
  whir::WhirCfg cfg;
};

struct whir_transcript {
  std::unique_ptr<whir::Transcript> t;
  whir::Duplex* duplex = nullptr;  // set for the built-in transcript
};

// Subset of ugu/util/io_util.h (include/ugu/util/io_util.h in UGU).
#pragma once

#include <string>

#include "ugu/image.h"

namespace ugu {

// WriteFaceIdAsText (src/util/io_util.cc:22-33): one face id per line in row
// order.
void WriteFaceIdAsText(const Image1i& face_id, const std::string& path);

}  // namespace ugu

// One-time upload of the constant tables to the context's device.
#include "gpb_pipeline.hpp"

namespace gpb {

void upload_tables(Context& c) {
  if (c.tables_ready) return;
  upload_schmid(c);
  c.tables_ready = true;
}

}  // namespace gpb

// Internal declarations shared by the npbcg translation units.
#ifndef NPBCG_INTERNAL_H
#define NPBCG_INTERNAL_H

#include <string>

namespace npbcg_detail {

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const std::string &msg);
void clear_error();

// Worker threads of the generator used by npbcg_makea calls from the calling
// thread (0: the default, NPBCG_GEN_THREADS or up to 16).
void set_generator_threads(int threads);

}  // namespace npbcg_detail

#endif

// Internal declarations shared by the npbcg translation units.
#ifndef NPBCG_INTERNAL_H
#define NPBCG_INTERNAL_H

#include <string>

namespace npbcg_detail {

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const std::string &msg);
void clear_error();

}  // namespace npbcg_detail

#endif

#pragma once

#include "../../../include/path_amd.h"
#include "../core/shading.h"
#include "../core/work_division.h"

#include <string>

namespace pt {
namespace host {

void default_work_division(pt_work_division_settings *w);
void default_settings(pt_settings *s);
IntegratorSettings integrator_settings(const pt_settings &s);
WorkDivisionSettings work_division_settings(const pt_work_division_settings &w);
bool load_settings_yaml(const std::string &path, pt_settings *s, std::string *err);
std::string settings_to_yaml(const pt_settings &s);

} // namespace host
} // namespace pt

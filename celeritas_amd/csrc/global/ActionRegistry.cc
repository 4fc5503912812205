//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/ActionRegistry.cc
//---------------------------------------------------------------------------//
#include "ActionInterface.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
void ActionRegistry::insert_impl(SPConstAction action)
{
    CELER_VALIDATE(action, << "cannot insert a null action");
    std::string label(action->label());
    CELER_VALIDATE(!label.empty(), << "action label is empty");
    CELER_VALIDATE(action->action_id() == this->next_id(),
                   << "incorrect action id {" << action->action_id()
                   << "} for action '" << label << "' (should be {"
                   << this->next_id() << "})");
    CELER_VALIDATE(!label_to_id_.count(label),
                   << "duplicate label '" << label << "' for action {"
                   << action->action_id() << "}");
    CELER_VALIDATE(this->next_id() < invalid_id8,
                   << "too many actions (at most 254 are addressable per slot)");
    label_to_id_[label] = action->action_id();
    labels_.push_back(std::move(label));
    actions_.push_back(std::move(action));
}

void ActionRegistry::insert(SPConstAction action)
{
    this->insert_impl(std::move(action));
}

void ActionRegistry::insert_mutable(std::shared_ptr<ActionInterface> action)
{
    this->insert_impl(action);
    mutable_.push_back(std::move(action));
}

auto ActionRegistry::action(ActionId id) const -> SPConstAction const&
{
    CELER_VALIDATE(id < actions_.size(), << "invalid action id " << id);
    return actions_[id];
}

ActionId ActionRegistry::find_action(std::string const& label) const
{
    auto iter = label_to_id_.find(label);
    if (iter == label_to_id_.end())
        return invalid_id;
    return iter->second;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd

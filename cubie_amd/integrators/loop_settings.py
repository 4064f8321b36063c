"""Loop setting names recognised by the solver (reference ode_loop.py:65-88),
plus ``save_times`` (an explicit table of save times, this engine's own)."""

ALL_LOOP_SETTINGS = {
    "save_every", "save_times", "summarise_every", "sample_summaries_every", "dt", "is_adaptive",
    "save_last", "save_regularly", "summarise_regularly", "state_location",
    "proposed_state_location", "parameters_location", "drivers_location",
    "proposed_drivers_location", "observables_location", "proposed_observables_location",
    "error_location", "counters_location", "state_summary_location",
    "observable_summary_location", "dt_location", "accept_step_location",
    "proposed_counters_location",
}

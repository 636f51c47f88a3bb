// Chemistry oracle (Li-Dryer H2/air): filled in by the chemistry restatement.

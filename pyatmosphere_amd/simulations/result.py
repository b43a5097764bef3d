"""Result containers with CSV checkpoint/resume (simulations/result.py:5-43 of
pyAtmosphere): save writes one column per measure with '{:.3e}' floats, and a
Result built with an existing `save_path` resumes from it."""
import numpy as np


class Result:
    save_float_format = "{:.3e}".format

    def __init__(self, channel, measures, max_size=None, save_path: str = ""):
        self.channel = channel
        self.measures = measures
        self.set_max_size(max_size)
        self.save_path = save_path
        if self.save_path:
            try:
                self.load_output()
                print(f"Loaded measures from {self.save_path}")
            except FileNotFoundError:
                pass

    def set_max_size(self, max_size):
        for measures in self.measures:
            measures.max_size = max_size

    def print_output(self):
        print(f"Len of the first measures: {len(self.measures[0])}")

    def plot_output(self):
        self.print_output()

    def as_df(self):
        import pandas as pd
        df = pd.DataFrame([measures.data for measures in self.measures]).T
        df.columns = [m.name for m in self.measures]
        return df

    def save_output(self):
        if not self.save_path:
            return
        self.as_df().to_csv(self.save_path, index=False, float_format=self.save_float_format)

    def load_output(self):
        import pandas as pd
        for measures, data in zip(self.measures, pd.read_csv(self.save_path).T.values):
            measures.data = [v for v in data.tolist() if not (isinstance(v, float) and np.isnan(v))]

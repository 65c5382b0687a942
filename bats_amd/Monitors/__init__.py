from .AbstractMonitor import AbstractMonitor
from .monitors import (AccuracyMonitor, LossMonitor, MonitorsManager, SilentLabelsMonitor,
                       SilentNeuronsMonitor, SpikeCountMonitor, TimeMonitor, ValueMonitor,
                       WeightsNormMonitor)
